// Building blocks of the fused per-window kernels (fused.hip): one workgroup
// of NT threads per window keeps the window's chips, spectra and surfaces in
// LDS; every FFT transforms all rows (or columns) of a chip at once in place.
//
// Correlations use row FFTs of length L >= search width plus a direct sum over
// template rows per frequency bin:
//   c[u,v] = (1/L) IFFT_L( sum_i conj(T_i) S_{i+u} )[v]
// which equals the reference's zero-padded 2-D FFT correlation
// (cuCorrFrequency.cu:63-79) at every valid lag.
#pragma once
#include <cfloat>
#include "cx.h"
#include "fft_inplace.h"

namespace ampcor {
namespace fused {

constexpr int NT = 512;   // threads per workgroup (8 wave64)

using cf = cx<float>;

__device__ __forceinline__ float pixel_amp(const void *img, int mode, int64_t off) {
    if (mode == 2) return fabsf(static_cast<const float *>(img)[off]);
    const float2 z = static_cast<const float2 *>(img)[off];
    return fast_sqrt(z.x * z.x + z.y * z.y);
}

// one pixel as loaded (real images: (value, 0)); raw_amp gives what pixel_amp gives
__device__ __forceinline__ float2 pixel_raw(const void *img, int mode, int64_t off) {
    if (mode == 2) return {static_cast<const float *>(img)[off], 0.f};
    return static_cast<const float2 *>(img)[off];
}

__device__ __forceinline__ float raw_amp(float2 z, int mode) {
    return mode == 2 ? fabsf(z.x) : fast_sqrt(z.x * z.x + z.y * z.y);
}

__device__ __forceinline__ cf pixel_cx(const void *img, int mode, int64_t off) {
    if (mode == 2) return {static_cast<const float *>(img)[off], 0.f};
    const float2 z = static_cast<const float2 *>(img)[off];
    if (mode == 1) return {fast_sqrt(z.x * z.x + z.y * z.y), 0.f};
    return {z.x, z.y};
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2. Giving every XCD a contiguous run of windows
// keeps neighbouring windows, whose chips overlap in the image, in one L2.
__device__ __forceinline__ int xcd_window(int b, int n) {
    const int x = b & 7, j = b >> 3;
    const int q = n >> 3, r = n & 7;     // XCD x runs q + (x < r) of the n blocks
    return x * q + min(x, r) + j;
}

// FP64 workgroup sum over NTH threads; `sh` holds NTH/64 doubles
template <int NTH = NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NTH / 64; ++i) s += sh[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// FP64 window sums -> nrm[u][v] = 1/sqrt((s2 - s1^2/n) t2 + FLT_EPSILON)
// (cuCorrNormalization.cu:230-314, cuCorrNormalizationSAT.cu:237-240).
// Vertical sums of s and s^2 over the first TH rows of every column come from
// NB row blocks summed by all threads at once (`part`: 2*NB*SW doubles of
// scratch); each wave then slides them down to its first lag row, and walks
// its contiguous lag rows with the horizontal window sums taken as differences
// of wave prefix scans (lanes own columns, CPL per lane; the scans run on DPP
// row shifts and row broadcasts, so they are VALU work with no LDS round trip).
// One barrier, which every thread of the workgroup must reach.
// ---------------------------------------------------------------------------
constexpr int NORM_BLOCKS = 4;

// DPP move of an FP64 value (both halves); lanes the row mask leaves out keep 0
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWMASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// DPP row shift of an FP64 value with every lane written: lanes whose source
// lies before the start of their row read 0 (bound_ctrl), so no zeroed
// destination is needed
template <int CTRL>
__device__ __forceinline__ double dpp_shr_f64(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// inclusive wave64 prefix sum: row_shr 1, 2, 4, 8 within rows of 16 lanes, then
// row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3
__device__ __forceinline__ double wave_scan_f64(double x) {
    x += dpp_shr_f64<0x111>(x);
    x += dpp_shr_f64<0x112>(x);
    x += dpp_shr_f64<0x114>(x);
    x += dpp_shr_f64<0x118>(x);
    x += dpp_f64<0x142, 0xa>(x);
    x += dpp_f64<0x143, 0xc>(x);
    return x;
}

__device__ __forceinline__ double lane63_f64(double x) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 63),
                            __builtin_amdgcn_readlane(__double2loint(x), 63));
}

// Lag-row phase of window_norms: part holds, for NB row blocks, the column sums
// of s ([b][SW]) then of s^2 ([NB + b][SW]). With blk = 0 the blocks cover rows
// [0, TH) and every wave slides its window down from row 0; with blk > 0 they
// are blocks of blk rows from row 0 down the whole search (TH a multiple of
// blk), and a wave starts from the TH / blk blocks at or above its first lag
// row, sliding at most blk - 1 rows. amp(r, c) is read for the slides only.
//
// Horizontal window sums of the column sums g (lane l holds columns l + 64m):
// * TWB = 0 (any TW): differences of the wave prefix over all SW columns,
//   read across lanes;
// * TWB = TW / 64 > 0: window(v) = sum_{c<TW} g[c] + sum_{j<v} (g[j+TW] - g[j]),
//   and column j + TW sits in the same lane as column j, so each 64-column
//   block of lag columns costs one wave scan of the differences plus one wave
//   sum of the first TW columns per lag row, with no value read across lanes.
template <int CPL, int NTH, int TWB = 0, class Amp>
__device__ __forceinline__ void window_norms_rows(Amp amp, const double *part, int NB, int SW, int TH, int TW,
                                                  int NLh, int NLw, double t2, float *nrm, int blk = 0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NW = NTH / 64;
    const int per = (NLh + NW - 1) / NW;
    const int u_beg = wave * per, u_end = min(NLh, u_beg + per);
    if (u_beg >= u_end) return;
    const int b_lo = blk ? u_beg / blk : 0, nb = blk ? TH / blk : NB;
    double g1[CPL], g2[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) {
        g1[m] = 0.0; g2[m] = 0.0;
        const int c = lane + 64 * m;
        if (c < SW) {
            for (int b = b_lo; b < b_lo + nb; ++b) {
                g1[m] += part[b * SW + c];
                g2[m] += part[(NB + b) * SW + c];
            }
            // slide down to u_beg (two independent chains)
            double d1 = 0.0, d2 = 0.0, e1 = 0.0, e2 = 0.0;
            int u = b_lo * blk;
#pragma unroll 2
            for (; u + 1 < u_beg; u += 2) {
                const double vn = amp(u + TH, c), vo = amp(u, c);
                const double wn = amp(u + 1 + TH, c), wo = amp(u + 1, c);
                d1 += vn - vo; d2 += vn * vn - vo * vo;
                e1 += wn - wo; e2 += wn * wn - wo * wo;
            }
            if (u < u_beg) {
                const double vn = amp(u + TH, c), vo = amp(u, c);
                d1 += vn - vo; d2 += vn * vn - vo * vo;
            }
            g1[m] += d1 + e1;
            g2[m] += d2 + e2;
        }
    }
    const double inv_n = 1.0 / double(TH * TW);
    #pragma unroll 1
    for (int u = u_beg; u < u_end; ++u) {
        if constexpr (TWB > 0) {
            static_assert(TWB < CPL, "lanes hold a column beyond the first TW");
            double e1 = g1[0], e2 = g2[0];
#pragma unroll
            for (int m = 1; m < TWB; ++m) { e1 += g1[m]; e2 += g2[m]; }
            double c1 = lane63_f64(wave_scan_f64(e1)), c2 = lane63_f64(wave_scan_f64(e2));
#pragma unroll
            for (int k = 0; k + TWB <= CPL; ++k) {
                if (64 * k >= NLw) break;
                const double h1 = k + TWB < CPL ? g1[k + TWB] : 0.0, h2 = k + TWB < CPL ? g2[k + TWB] : 0.0;
                const double d1 = h1 - g1[k], d2 = h2 - g2[k];
                const double x1 = wave_scan_f64(d1), x2 = wave_scan_f64(d2);
                const int v = lane + 64 * k;
                if (v < NLw) {
                    const double s1 = c1 + (x1 - d1), s2 = c2 + (x2 - d2);
                    const double norm2 = (s2 - s1 * s1 * inv_n) * t2;
                    nrm[u * NLw + v] = __builtin_amdgcn_rsqf(float(norm2 + double(FLT_EPSILON)));
                }
                c1 += lane63_f64(x1);
                c2 += lane63_f64(x2);
            }
        } else {
        double p1[CPL], p2[CPL];
        double c1 = 0.0, c2 = 0.0;
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
            const double x1 = wave_scan_f64(g1[m]) + c1, x2 = wave_scan_f64(g2[m]) + c2;
            p1[m] = x1; p2[m] = x2;
            c1 = lane63_f64(x1);
            c2 = lane63_f64(x2);
        }
        #pragma unroll 1
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
                // the cancellation is resolved in FP64; the reciprocal square root of
                // the FP32 result is the hardware one (within 1 ulp)
                nrm[u * NLw + v] = __builtin_amdgcn_rsqf(float(norm2 + double(FLT_EPSILON)));
            }
        }
        }
        if (u + 1 < u_end) {
#pragma unroll
            for (int m = 0; m < CPL; ++m) {
                const int c = lane + 64 * m;
                if (c < SW) {
                    const double vn = amp(u + TH, c), vo = amp(u, c);
                    g1[m] += vn - vo;
                    g2[m] += vn * vn - vo * vo;
                }
            }
        }
    }
}

// Window sums from column sums of row blocks covering the whole search height
// (SH rows, blocks of NORM_BLK rows; part holds 2 * ceil(SH / NORM_BLK) * SW
// doubles), so that no wave slides far to its first lag row. TH must be a
// multiple of NORM_BLK.
constexpr int NORM_BLK = 16;
template <int CPL, int NTH = NT, int TWB = 0, class Amp>
__device__ __forceinline__ void window_norms_blk(Amp amp, int SH, int SW, int TH, int TW, int NLh, int NLw,
                                                 double t2, float *nrm, double *part) {
    const int NB = (SH + NORM_BLK - 1) / NORM_BLK;
    #pragma unroll 1
    for (int t = threadIdx.x; t < NB * SW; t += NTH) {
        const int b = t / SW, c = t - b * SW;
        const int r0 = b * NORM_BLK, r1 = min(SH, r0 + NORM_BLK);
        double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
        int r = r0;
        for (; r + 1 < r1; r += 2) {
            const double x = amp(r, c), y = amp(r + 1, c);
            a1 += x; a2 += x * x;
            b1 += y; b2 += y * y;
        }
        if (r < r1) {
            const double x = amp(r, c);
            a1 += x; a2 += x * x;
        }
        part[b * SW + c] = a1 + b1;
        part[(NB + b) * SW + c] = a2 + b2;
    }
    __syncthreads();
    window_norms_rows<CPL, NTH, TWB>(amp, part, NB, SW, TH, TW, NLh, NLw, t2, nrm, NORM_BLK);
}

template <int CPL, int NTH = NT, int TWB = 0, class Amp>
__device__ __forceinline__ void window_norms(Amp amp, int SW, int TH, int TW, int NLh, int NLw, double t2, float *nrm,
                             double *part) {
    constexpr int NB = NORM_BLOCKS;
    const int rows_per = (TH + NB - 1) / NB;
    #pragma unroll 1
    for (int t = threadIdx.x; t < NB * SW; t += NTH) {
        const int b = t / SW, c = t - b * SW;
        const int r0 = b * rows_per, r1 = min(TH, r0 + rows_per);
        double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
        int r = r0;
        for (; r + 1 < r1; r += 2) {
            const double x = amp(r, c), y = amp(r + 1, c);
            a1 += x; a2 += x * x;
            b1 += y; b2 += y * y;
        }
        if (r < r1) {
            const double x = amp(r, c);
            a1 += x; a2 += x * x;
        }
        part[b * SW + c] = a1 + b1;
        part[(NB + b) * SW + c] = a2 + b2;
    }
    __syncthreads();
    window_norms_rows<CPL, NTH, TWB>(amp, part, NB, SW, TH, TW, NLh, NLw, t2, nrm);
}

// ---------------------------------------------------------------------------
// In-place split of packed real-pair spectra. Signal j (pitch P) holds
// Z = FFT(a + i b) in digit-reversed order (fft_nat2dr); afterwards slots
// [0, H) hold A_k (row 2j) and [H, 2H) hold B_k (row 2j+1), k < H = L/2, with
// slot 0 = (DC, Nyquist) of that row. A round covers whole signals only: the
// reads of signal j (bins k and L-k) must all happen before any write to j.
// ---------------------------------------------------------------------------
template <int R1, int R2, int NTH = NT>
__device__ __forceinline__ void split_pairs_inplace(cf *buf, int nsig, int P) {
    constexpr int L = R1 * R2, H = L / 2, U = 4;
    // each thread keeps one bin for the whole call (its storage indices computed
    // once): bins 1..H-1 on the first SPR (H-1) threads, bin 0 on the next SPR,
    // so only the wave that holds the bin-0 threads takes both paths; a round
    // covers U SPR whole signals
    constexpr int SPR = NTH / H, NG = SPR * (H - 1);
    static_assert(SPR >= 1, "a pass must hold at least one signal");
    const int t = threadIdx.x;
    const bool gen = t < NG, act = t < NG + SPR;
    const int jj = gen ? t / (H - 1) : t - NG;
    const int k = gen ? 1 + t % (H - 1) : 0;
    const int pk = dr_pos<R1, R2>(k), pc = dr_pos<R1, R2>(gen ? L - k : H);
    #pragma unroll 1
    for (int j0 = 0; j0 < nsig; j0 += U * SPR) {
        cf av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + jj + u * SPR;
            if (act && j < nsig) {
                const cf *z = buf + j * P;
                const cf zk = z[pk], zp = z[pc];
                if (gen) {
                    const cf zc = conj(zp);
                    av[u] = (zk + zc) * 0.5f;
                    const cf d = (zk - zc) * 0.5f;
                    bv[u] = {d.y, -d.x};
                } else {
                    av[u] = {zk.x, zp.x};
                    bv[u] = {zk.y, zp.y};
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + jj + u * SPR;
            if (act && j < nsig) {
                buf[j * P + k] = av[u];
                buf[j * P + H + k] = bv[u];
            }
        }
        __syncthreads();
    }
}

// address of row r's half spectrum after split_pairs_inplace
template <int H>
__device__ __forceinline__ int row_slot(int r, int P) { return (r >> 1) * P + (r & 1) * H; }

// ---------------------------------------------------------------------------
// MAC over template rows [i_beg, i_end): q[k] += conj(T_i[f]) S_{i+u0+k}[f]
// (slot 0 packs two real bins: (DC, Nyquist) multiply component-wise; with
// DC = false every lane takes the complex form and the caller computes slot 0's
// sums elsewhere: dc_lag_sum). Rows sit in pairs (row r at slot (r >> 1) P +
// (r & 1) H); i_beg and t_off are even and the range a multiple of 8, so a
// thread's search rows keep the parity of u0 and every load of a block is a
// fixed offset from one of two addresses. Template row i meets search row
// i + u0 + k, so a search row at or past SH meets only lags u0 + k >=
// SH - (i_end - 1), past the last lag the callers keep (SH - template height + 1
// lags), and the callers' LDS continues past their search rows: such rows are
// read without a bound check and only reach sums that are discarded.
// ---------------------------------------------------------------------------
template <int H, int K, bool DC = true>
__device__ __forceinline__ void mac_rows(cf (&q)[K], const cf *S, int Sp, const cf *Tsp, int Tp, int t_off,
                                         int i_beg, int i_end, int u0, int SH, int f) {
    const bool dc = DC && f == 0;
    const int p = u0 & 1;
    const int e0 = p ? H : 0, e1 = p ? Sp : H;      // slot offset of even / odd block rows
    (void)SH;
    #pragma unroll 1
    for (int i0 = i_beg; i0 < i_end; i0 += 8) {
        const cf *sb = S + ((i0 + u0) >> 1) * Sp + f;
        const cf *tb = Tsp + ((i0 - t_off) >> 1) * Tp + f;
        cf sw[K + 7];
#pragma unroll
        for (int m = 0; m < K + 7; ++m) sw[m] = sb[(m >> 1) * Sp + ((m & 1) ? e1 : e0)];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const cf tv = tb[(i >> 1) * Tp + (i & 1) * H];
            // one branch-free form for every lane: conj(t) s, or for slot 0 the
            // component-wise product (t.x s.x, t.y s.y)
            const float c0 = tv.x, c1 = dc ? 0.f : tv.y, c2 = dc ? tv.y : tv.x, c3 = dc ? 0.f : -tv.y;
            // two chained FMAs per component (the accumulator is the addend)
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const cf s = sw[i + k];
                q[k].x = fmaf(c1, s.y, fmaf(c0, s.x, q[k].x));
                q[k].y = fmaf(c3, s.x, fmaf(c2, s.y, q[k].y));
            }
        }
    }
}

// Slot 0 of mac_rows for one lag u and one component (0: DC, 1: Nyquist):
// acc += T_i[0].comp * S_{s_off+i+u}[0].comp over template rows i in [0, n) of Tsp,
// in mac_rows' order, so for the lags a caller keeps (whose search rows all lie
// below SH; rows at or past it read as zero here) the sums agree bit for bit
// with mac_rows' slot-0 lane
template <int H, int UNROLL = 8>
__device__ __forceinline__ float dc_lag_sum(float acc, const cf *S, int Sp, const cf *Tsp, int Tp, int n, int s_off,
                                            int u, int SH, int comp) {
    const float *s = reinterpret_cast<const float *>(S) + comp, *t = reinterpret_cast<const float *>(Tsp) + comp;
    #pragma unroll UNROLL
    for (int i = 0; i < n; ++i) {
        const int r = s_off + i + u;
        const float sv = (r < SH) ? s[2 * row_slot<H>(r, Sp)] : 0.f;
        acc = fmaf(t[2 * row_slot<H>(i, Tp)], sv, acc);
    }
    return acc;
}

// Inverse of the accumulated Q[u][f] (Qs: [NLh][H]) into corr[u][v] * nrm[u][v]
// for v < NLw: lag pairs (2p, 2p+1) share one complex transform placed in Y
// (npair signals, pitch P, digit-reversed input order).
template <int R1, int R2, int NTH = NT>
__device__ __forceinline__ void inverse_lags(const cf *Qs, int NLh, int NLw, cf *Y, int P, const cx<float> *tw, const float *nrm,
                             float *corr) {
    constexpr int L = R1 * R2, H = L / 2;
    const int npair = (NLh + 1) / 2;
    for (int i = threadIdx.x; i < npair * L; i += NTH) {
        const int p = i / L, k = i - p * L;
        const int ua = 2 * p, ub = ua + 1;
        const cf *qa = Qs + ua * H;
        const bool hb = ub < NLh;
        const cf *qb = Qs + (hb ? ub : ua) * H;
        cf y;
        if (k == 0) y = {qa[0].x, hb ? qb[0].x : 0.f};
        else if (k == H) y = {qa[0].y, hb ? qb[0].y : 0.f};
        else if (k < H) {
            const cf a = qa[k], b = hb ? qb[k] : cf{0.f, 0.f};
            y = {a.x - b.y, a.y + b.x};
        } else {
            const cf a = qa[L - k], b = hb ? qb[L - k] : cf{0.f, 0.f};
            y = {a.x + b.y, b.x - a.y};
        }
        Y[p * P + dr_pos<R1, R2>(k)] = y;
    }
    __syncthreads();
    fft_dr2nat<float, +1, R1, R2, NTH>(Y, npair, P, 1, tw);
    const float invL = 1.0f / float(L);
    for (int i = threadIdx.x; i < NLh * NLw; i += NTH) {
        const int u = i / NLw, v = i - u * NLw;
        const cf y = Y[(u >> 1) * P + v];
        corr[i] = ((u & 1) ? y.y : y.x) * invL * nrm[i];
    }
    __syncthreads();
}

// argmax over n values in LDS with the reference tie rule
__device__ __forceinline__ Top2<float> block_argmax(const float *v, int n, double *red) {
    Top2<float> t;
    for (int i = threadIdx.x; i < n; i += NT) t.add(v[i], maxloc_key(uint32_t(i)));
    float *shv = reinterpret_cast<float *>(red + 8);
    float *shv2 = shv + 8;
    uint64_t *shk = reinterpret_cast<uint64_t *>(red + 16);
    t.reduce<NT>(shv, shv2, shk);
    return t;
}

}  // namespace fused
}  // namespace ampcor
