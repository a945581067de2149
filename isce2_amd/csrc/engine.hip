// Engine host code: parameters, FFT plans, device buffers, the staged pipeline
// and the C ABI declared in include/isce2_amd.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include "engine.h"
#include "stage_kernels.h"

namespace ampcor {

// ---------------------------------------------------------------------------
void DevBuf::reserve(size_t n) {
    if (n <= bytes) return;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t want = std::max<size_t>(n, 256);
    if (hipMalloc(&ptr, want) != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        throw Error(AMPCOR_E_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
    }
    bytes = want;
}
DevBuf::~DevBuf() {
    if (ptr) (void)hipFree(ptr);
}

hipEvent_t Engine::take_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    return e;
}

void Engine::prof_collect() {
    for (auto &r : prof_pending) {
        HIP_OK(hipEventSynchronize(r.b));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
        auto &acc = prof_acc[r.name];
        acc.launches += 1;
        acc.ms += ms;
        acc.windows += r.windows;
        event_pool.push_back(r.a);
        event_pool.push_back(r.b);
    }
    prof_pending.clear();
}

Engine::~Engine() {
    for (auto &r : prof_pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : event_pool) (void)hipEventDestroy(e);
}

FftDesc make_schedule(int n) {
    FftDesc d{};
    d.n = n;
    int m = n;
    const int pow2[] = {16, 8, 4, 2};
    const int odd[] = {3, 5, 7, 11, 13, 17};
    for (int r : pow2)
        while (m % r == 0 && m > 1) { d.radix[d.npass++] = r; m /= r; }
    for (int r : odd)
        while (m % r == 0) { d.radix[d.npass++] = r; m /= r; }
    if (m != 1 || d.npass > 14) {
        d.npass = 0;
        d.direct = 1;
    }
    return d;
}

const FftPlan &Engine::plan(int n) {
    auto it = plans.find(n);
    if (it != plans.end()) return *it->second;
    forbid_in_nosync("build an FFT plan");
    std::unique_ptr<FftPlan> p(new FftPlan());
    p->n = n;
    p->desc32 = p->desc64 = make_schedule(n);
    std::vector<float> t32(2 * n);
    std::vector<double> t64(2 * n);
    for (int k = 0; k < n; ++k) {
        // e^{-2 pi i k/n} in FP64; sin/cos of reduced angles for accuracy
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * k / n;
        const double c = double(cosl(a)), s = double(sinl(a));
        t64[2 * k] = c; t64[2 * k + 1] = s;
        t32[2 * k] = float(c); t32[2 * k + 1] = float(s);
    }
    p->tw32.reserve(t32.size() * sizeof(float));
    p->tw64.reserve(t64.size() * sizeof(double));
    HIP_OK(hipMemcpy(p->tw32.ptr, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(p->tw64.ptr, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
    p->desc32.tw = p->tw32.ptr;
    p->desc64.tw = p->tw64.ptr;
    auto &ref = *p;
    plans[n] = std::move(p);
    return ref;
}

// cuAmpcorParameter::setupParameters (cuAmpcorParameter.cpp:70-98), with each
// axis sized by its own half search range (the reference swaps them, :90-91).
void Engine::setup(const ampcor_params &q) {
    if (q.window_h <= 0 || q.window_w <= 0) throw Error(AMPCOR_E_ARG, "window size must be positive");
    if (q.hs_down <= 0 || q.hs_across <= 0) throw Error(AMPCOR_E_ARG, "half search range must be positive");
    if (q.raw_ovs <= 0 || q.surf_ovs <= 0) throw Error(AMPCOR_E_ARG, "oversampling factors must be positive");
    if (q.zoom <= 0) throw Error(AMPCOR_E_ARG, "corrSurfaceZoomInWindow must be positive");
    if (q.corr_stat <= 0) throw Error(AMPCOR_E_ARG, "corrStatWindowSize must be positive");
    ampcor_sizes s{};
    s.zoom = std::min(q.zoom, std::min(q.hs_across, q.hs_down) * 2 * q.raw_ovs);
    s.half_zoom = s.zoom / (2 * q.raw_ovs);
    s.corr_raw_zoom_h = std::min(q.corr_stat, 2 * q.hs_down + 1);
    s.corr_raw_zoom_w = std::min(q.corr_stat, 2 * q.hs_across + 1);
    s.search_h = q.window_h + 2 * q.hs_down;
    s.search_w = q.window_w + 2 * q.hs_across;
    s.zoom_in_h = q.window_h + 2 * s.half_zoom;
    s.zoom_in_w = q.window_w + 2 * s.half_zoom;
    s.os_win_h = q.window_h * q.raw_ovs;
    s.os_win_w = q.window_w * q.raw_ovs;
    s.os_search_h = s.zoom_in_h * q.raw_ovs;
    s.os_search_w = s.zoom_in_w * q.raw_ovs;
    s.raw_corr_h = s.search_h - q.window_h + 1;
    s.raw_corr_w = s.search_w - q.window_w + 1;
    s.os_corr_h = s.os_search_h - s.os_win_h + 1;
    s.os_corr_w = s.os_search_w - s.os_win_w + 1;
    s.surf_out = s.zoom * q.surf_ovs;
    if (s.os_corr_h - 1 != s.zoom || s.os_corr_w - 1 != s.zoom)
        throw Error(AMPCOR_E_ARG, "corrSurfaceZoomInWindow (" + std::to_string(q.zoom) +
                                      ") must be a multiple of 2*rawDataOversamplingFactor");
    if (s.half_zoom <= 0) throw Error(AMPCOR_E_ARG, "zoom-in window smaller than 2*rawDataOversamplingFactor");
    if (std::max({s.search_h, s.search_w, s.os_search_h, s.os_search_w, s.surf_out}) > 4096)
        throw Error(AMPCOR_E_UNSUPPORTED, "chip or surface side above 4096");
    s.fused = fused_supported(q, s) ? 1 : 0;
    // the correlation is linear for every transform length >= the search chip
    // (cuCorrFrequency.cu:16-43 uses the chip size): take the next 5-smooth length
    auto smooth = [](int n) {
        for (int m = n;; ++m) {
            int r = m;
            for (int f : {2, 3, 5})
                while (r % f == 0) r /= f;
            if (r == 1) return m;
        }
    };
    s.fft_corr_raw_h = smooth(s.search_h);
    s.fft_corr_raw_w = smooth(s.search_w);
    s.fft_corr_os_h = smooth(s.os_search_h);
    s.fft_corr_os_w = smooth(s.os_search_w);
    p = q;
    sz = s;
    configured = true;
    // the fused kernels' tables are built here, so that runs (the allocation-free
    // entry in particular) find them in place
    if (s.fused) {
        HIP_OK(hipSetDevice(device));
        (void)fused_tables();
    }
}

// Replace the radix schedule of the length-n staged FFTs (autotuning hook).
// Radices must be supported by the Stockham passes and multiply to n.
void Engine::set_fft_radices(int n, const int32_t *radices, int count) {
    if (n < 2 || count < 1 || count > 14 || !radices) throw Error(AMPCOR_E_ARG, "bad FFT schedule");
    long prod = 1;
    for (int i = 0; i < count; ++i) {
        const int r = radices[i];
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 11 && r != 13 && r != 16 && r != 17)
            throw Error(AMPCOR_E_ARG, "unsupported FFT radix " + std::to_string(r));
        prod *= r;
    }
    if (prod != n) throw Error(AMPCOR_E_ARG, "FFT radices do not multiply to " + std::to_string(n));
    FftPlan &p = const_cast<FftPlan &>(plan(n));
    for (FftDesc *d : {&p.desc32, &p.desc64}) {
        d->direct = 0;
        d->npass = count;
        for (int i = 0; i < count; ++i) d->radix[i] = radices[i];
    }
}

void Engine::ensure_sinc_table() {
    if (sinc_table.ptr) return;
    forbid_in_nosync("build the sinc table");
    // cuSincOverSamplerR2R table (cuSincOverSampler.cu:21-80, constants cuSincOverSampler.h:27-57)
    const double beta = 0.75, relfiltlen = 6.0, pedestal = 0.0, PI = 3.14159265359;
    sinc_len = int(relfiltlen / beta + 0.5);
    sinc_dec = 4096;
    const int ncoef = sinc_len * sinc_dec;
    const double wgthgt = (1.0 - pedestal) / 2.0, soff = (ncoef - 1.0) / 2.0;
    std::vector<float> t(ncoef + 1);
    for (int i = 0; i <= ncoef; ++i) {
        const double wa = i - soff;
        const double wgt = (1.0 - wgthgt) + wgthgt * std::cos(PI * wa / soff);
        const double x = wa * beta / sinc_dec * PI;
        const double f = x != 0.0 ? std::sin(x) / x : 1.0;
        t[i] = float(f * wgt);
    }
    sinc_table.reserve(t.size() * sizeof(float));
    HIP_OK(hipMemcpy(sinc_table.ptr, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------------------
// staged pipeline: launch helpers
// ---------------------------------------------------------------------------
namespace {
// elems: complex elements one FFT block holds (x2 buffers in LDS), the
// "fft_block_elems" launch option
template <class T, int DIR>
void fft_rows(const FftPlan &pl, cx<T> *data, int64_t nrows, int elems, hipStream_t st) {
    const FftDesc &d = std::is_same<T, float>::value ? pl.desc32 : pl.desc64;
    const int g = std::max(1, elems / pl.n);
    const size_t lds = 2 * size_t(g) * pl.n * sizeof(cx<T>);
    const int64_t nblk = (nrows + g - 1) / g;
    for (int64_t b0 = 0; b0 < nblk; b0 += 1 << 30) {
        const int64_t nb = std::min<int64_t>(nblk - b0, 1 << 30);
        hipLaunchKernelGGL((k_fft_rows<T, DIR>), dim3(unsigned(nb)), dim3(256), lds, st,
                           data + b0 * g * pl.n, int(std::min<int64_t>(nrows - b0 * g, nb * g)), g, d);
    }
}

// columns per block of a column FFT of length pl.n over w columns: from the
// element budget, or a fixed width (whole 64/128-byte column segments per row)
// when the LDS for it is available
template <class T>
int cols_per_block(const FftPlan &pl, int w, int elems, int width) {
    int c = std::max(1, std::min(w, std::min(16, elems / pl.n)));
    if (width > 0 && 2 * size_t(width) * pl.n * sizeof(cx<T>) <= 160 * 1024) c = std::min(w, width);
    return c;
}

template <class T, int DIR>
void fft_cols(const FftPlan &pl, cx<T> *data, int nimg, int w, int elems, hipStream_t st, int width = 0) {
    const FftDesc &d = std::is_same<T, float>::value ? pl.desc32 : pl.desc64;
    const int c = cols_per_block<T>(pl, w, elems, width);
    const size_t lds = 2 * size_t(c) * pl.n * sizeof(cx<T>);
    const int threads = c >= 16 ? 512 : 256;
    hipLaunchKernelGGL((k_fft_cols<T, DIR>), dim3(idivup(w, c), nimg), dim3(threads), lds, st, data, w, c, d);
}

template <class T, int DIR>
void fft_rows_sel(const FftPlan &pl, cx<T> *data, int nimg, int H, int sel, int split, int off1, int elems,
                  hipStream_t st) {
    const FftDesc &d = std::is_same<T, float>::value ? pl.desc32 : pl.desc64;
    const int g = std::max(1, elems / pl.n);
    const size_t lds = 2 * size_t(g) * pl.n * sizeof(cx<T>);
    hipLaunchKernelGGL((k_fft_rows_sel<T, DIR>), dim3(idivup(nimg * sel, g)), dim3(256), lds, st, data, nimg, H, sel,
                       split, off1, g, d);
}

template <class T>
int fft_elems(const Engine &e) { return sizeof(T) == 8 ? std::max(1024, e.fft_block_elems / 2) : e.fft_block_elems; }

// fixed column-block width for the FP32 column FFTs (the FP64 ones use the element budget)
template <class T>
int cols_width(const Engine &e) { return sizeof(T) == 4 ? e.fft_cols_width : 0; }

// Inverse 2D FFT of the quadrant padding of h x w spectra `in` into H x W,
// keeping the real image `out` only (OUT 1 |z|, OUT 2 Re z); `work` holds the
// H x W complex row-pass result, of which the all-zero rows are neither written
// nor read
template <class T, int DIR, int OUT>
void fft2_quad(Engine &e, const cx<T> *in, int nimg, int h, int w, cx<T> *work, int H, int W, T *out,
               hipStream_t st) {
    const FftPlan &pr = e.plan(W), &pc = e.plan(H);
    const bool f32 = std::is_same<T, float>::value;
    const int elems = fft_elems<T>(e), sel = 2 * (h / 2);
    const int g = std::max(1, elems / W);
    hipLaunchKernelGGL((k_fft_rows_quad<T, DIR>), dim3(idivup(nimg * sel, g)), dim3(256),
                       2 * size_t(g) * W * sizeof(cx<T>), st, in, nimg, h, w, work, H, g, f32 ? pr.desc32 : pr.desc64);
    const int c = cols_per_block<T>(pc, W, elems, cols_width<T>(e));
    hipLaunchKernelGGL((k_fft_cols_quad<T, DIR, OUT>), dim3(idivup(W, c), nimg), dim3(c >= 16 ? 512 : 256),
                       2 * size_t(c) * H * sizeof(cx<T>), st, work, W, c, h / 2, f32 ? pc.desc32 : pc.desc64, out);
}

// 2D FFT of which only rows [0, keep) of the result are used: columns first,
// then only those rows
template <class T, int DIR>
void fft2_rows_kept(Engine &e, cx<T> *data, int nimg, int h, int w, int keep, hipStream_t st) {
    fft_cols<T, DIR>(e.plan(h), data, nimg, w, fft_elems<T>(e), st, cols_width<T>(e));
    fft_rows_sel<T, DIR>(e.plan(w), data, nimg, h, keep, keep, 0, fft_elems<T>(e), st);
}

// 2D FFT of nimg images h x w (rows then columns), in place
template <class T, int DIR>
void fft2(Engine &e, cx<T> *data, int nimg, int h, int w, hipStream_t st) {
    // double elements take twice the LDS: half the element budget (fft_elems)
    fft_rows<T, DIR>(e.plan(w), data, int64_t(nimg) * h, fft_elems<T>(e), st);
    fft_cols<T, DIR>(e.plan(h), data, nimg, w, fft_elems<T>(e), st, cols_width<T>(e));
}

// forward 2D FFT of h x w chips read from [nimg][ih][iw] sources, cropped at
// `start` and deramped with `tab` when given (k_fft_rows_chip), into dst
template <class T, int DIR>
void fft2_chip(Engine &e, const cx<T> *src, int ih, int iw, const int32_t *start, const cx<double> *tab, cx<T> *dst,
               int nimg, int h, int w, hipStream_t st) {
    const FftPlan &pl = e.plan(w);
    const int g = std::max(1, fft_elems<T>(e) / w);
    hipLaunchKernelGGL((k_fft_rows_chip<T, DIR>), dim3(idivup(nimg * h, g)), dim3(256), 2 * size_t(g) * w * sizeof(cx<T>),
                       st, src, ih, iw, start, tab, dst, nimg, h, g, sizeof(T) == 4 ? pl.desc32 : pl.desc64);
    fft_cols<T, DIR>(e.plan(h), dst, nimg, w, fft_elems<T>(e), st, cols_width<T>(e));
}

// x2 oversampled amplitude `out` ([nimg][2h][2w]) of h x w chips read from
// [nimg][ih][iw] sources (cropped at `start`, deramped with `tab` when given),
// by the separable interpolation (k_os2_cols, then k_os2_rows); `Y` holds the
// [nimg][h][w] odd rows between the passes
// in-place kernels for the chip sides of the common windows (32, 64, 128, 256
// and their zoom chips, 8 pixels wider); false for other sides
template <class T>
const cx<T> *tw_of(Engine &e, int n) { return static_cast<const cx<T> *>(sizeof(T) == 4 ? e.plan(n).tw32.ptr : e.plan(n).tw64.ptr); }

template <class T, int R1, int R2, int CB = 16, int NTH = 256>
void os2_cols_launch(Engine &e, const cx<T> *src, int ih, int iw, const int32_t *start, const cx<double> *tab,
                     int nimg, int w, cx<T> *Y, hipStream_t st) {
    constexpr int N = R1 * R2;
    const size_t lds = (size_t(N) * (CB + 1) + N) * sizeof(cx<T>);
    hipLaunchKernelGGL((k_os2_cols_f<T, R1, R2, CB, NTH>), dim3(idivup(w, CB), nimg), dim3(NTH), lds, st, src, ih, iw,
                       start, tab, w, tw_of<T>(e, N), Y);
}

// output rows per block of k_os2_rows_f: 16 where the second factor is a long
// odd DFT (its pass has G * R1 work items), else 8 (measured on window-256 pairs)
template <class T, int R2>
constexpr int os2_rows_g() { return sizeof(T) == 4 && R2 >= 17 ? 16 : 8; }

// the same count by chip width (the widths os2_rows_fast runs: 136 = 8 x 17 is the
// one with a long second factor)
template <class T>
int os2_rows_per_block(int w) { return sizeof(T) == 4 && w == 136 ? 16 : 8; }

template <class T, int R1, int R2, int NTH = 256>
void os2_rows_launch(Engine &e, const cx<T> *src, int ih, int iw, const int32_t *start, const cx<double> *tab,
                     const cx<T> *Y, int nimg, int h, T *out, double *part, hipStream_t st) {
    constexpr int N = R1 * R2, G = os2_rows_g<T, R2>();
    // k_mean_parts reads one block sum per G rows: the width table must agree
    if (G != os2_rows_per_block<T>(N)) throw Error(AMPCOR_E_ARG, "row block size table out of step");
    const size_t lds = (size_t(G) * (N + 1) + N) * sizeof(cx<T>) + size_t(G) * N * sizeof(T);
    hipLaunchKernelGGL((k_os2_rows_f<T, R1, R2, G, NTH>), dim3(idivup(nimg * 2 * h, G)), dim3(NTH), lds, st, src, ih, iw,
                       start, tab, Y, nimg, h, tw_of<T>(e, N), out, part);
}

template <class T>
bool os2_cols_fast(Engine &e, const cx<T> *src, int ih, int iw, const int32_t *start, const cx<double> *tab,
                   int nimg, int h, int w, cx<T> *Y, hipStream_t st) {
    switch (h) {
        case 32: os2_cols_launch<T, 8, 4>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 40: os2_cols_launch<T, 8, 5>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 64: os2_cols_launch<T, 8, 8>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 72: os2_cols_launch<T, 8, 9>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 128: os2_cols_launch<T, 16, 8>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 136: os2_cols_launch<T, 8, 17>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 256: os2_cols_launch<T, 16, 16>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        case 264: os2_cols_launch<T, 11, 24>(e, src, ih, iw, start, tab, nimg, w, Y, st); return true;
        default: return false;
    }
}

// chip widths with an in-place row kernel
bool os2_rows_has(int w) {
    switch (w) {
        case 32: case 40: case 64: case 72: case 128: case 136: case 256: case 264: return true;
        default: return false;
    }
}

template <class T>
bool os2_rows_fast(Engine &e, const cx<T> *src, int ih, int iw, const int32_t *start, const cx<double> *tab,
                   const cx<T> *Y, int nimg, int h, int w, T *out, double *part, hipStream_t st) {
    switch (w) {
        case 32: os2_rows_launch<T, 8, 4>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 40: os2_rows_launch<T, 8, 5>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 64: os2_rows_launch<T, 8, 8>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 72: os2_rows_launch<T, 8, 9>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 128: os2_rows_launch<T, 16, 8>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 136: os2_rows_launch<T, 8, 17>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 256: os2_rows_launch<T, 16, 16>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        case 264: os2_rows_launch<T, 24, 11, 192>(e, src, ih, iw, start, tab, Y, nimg, h, out, part, st); return true;
        default: return false;
    }
}

// `part`: block sums for k_mean_parts (in-place row kernel only; null for none)
template <class T>
void os2_separable(Engine &e, const cx<T> *src, int ih, int iw, const int32_t *start, const cx<double> *tab, int nimg,
                   int h, int w, cx<T> *Y, T *out, hipStream_t st, double *part = nullptr) {
    const bool f32 = std::is_same<T, float>::value;
    const FftPlan &pc = e.plan(h), &pr = e.plan(w);
    const int elems = fft_elems<T>(e);
    if (e.oversampler == 2) {
        const bool cols = os2_cols_fast<T>(e, src, ih, iw, start, tab, nimg, h, w, Y, st);
        if (cols && os2_rows_fast<T>(e, src, ih, iw, start, tab, Y, nimg, h, w, out, part, st)) return;
        if (cols) {
            const int g = std::max(1, std::min(elems / w, (81920 - w * int(sizeof(cx<T>))) /
                                                              (w * int(2 * sizeof(cx<T>) + sizeof(T)))));
            hipLaunchKernelGGL(k_os2_rows<T>, dim3(idivup(nimg * 2 * h, g)), dim3(256),
                               (2 * size_t(g) + 1) * w * sizeof(cx<T>) + size_t(g) * w * sizeof(T), st, src, ih, iw,
                               start, tab, Y, nimg, h, g, f32 ? pr.desc32 : pr.desc64, out);
            return;
        }
    }
    // up to 16 columns per block (whole 128-byte row segments) within 80 KB of LDS
    const int cb = std::max(1, std::min(w, std::min(16, (81920 / int(h * sizeof(cx<T>)) - 1) / 2)));
    hipLaunchKernelGGL(k_os2_cols<T>, dim3(idivup(w, cb), nimg), dim3(256), (2 * size_t(cb) + 1) * h * sizeof(cx<T>),
                       st, src, ih, iw, start, tab, w, cb, f32 ? pc.desc32 : pc.desc64, Y);
    // rows per block from the element budget, within 80 KB of LDS
    const int row_bytes = w * int(2 * sizeof(cx<T>) + sizeof(T));
    const int g = std::max(1, std::min(elems / w, (81920 - w * int(sizeof(cx<T>))) / row_bytes));
    hipLaunchKernelGGL(k_os2_rows<T>, dim3(idivup(nimg * 2 * h, g)), dim3(256),
                       (2 * size_t(g) + 1) * w * sizeof(cx<T>) + size_t(g) * w * sizeof(T), st, src, ih, iw, start, tab,
                       Y, nimg, h, g, f32 ? pr.desc32 : pr.desc64, out);
}

// Row-spectrum transform lengths with an in-place kernel (k_rowspec_f, FP32):
// L = R1 * R2, listed by L. The correlation is the same for any L of at least
// the search width; these cover the uneven windows' search widths with little
// padding and prime factors <= 17 (which the lag-row inverse needs).
struct RowLen { int L, R1, R2; };
constexpr RowLen kRowLens[] = {{72, 8, 9},    {80, 16, 5},   {104, 8, 13},  {128, 16, 8},  {144, 16, 9},
                               {168, 8, 21},  {195, 13, 15}, {256, 16, 16}, {272, 16, 17}, {289, 17, 17},
                               {325, 13, 25}, {528, 16, 33}};

// smallest listed length in [sw, sw + 48], or 0
int row_len_fast(int sw) {
    for (const RowLen &r : kRowLens)
        if (r.L >= sw && r.L <= sw + 48) return r.L;
    return 0;
}

template <int R1, int R2, int G = 8, int NTH = 256>
void rowspec_launch(Engine &e, const float *img, int nimg, int h, int w, const double *mean, cx<float> *spec,
                    hipStream_t st) {
    constexpr int L = R1 * R2;
    hipLaunchKernelGGL((k_rowspec_f<float, R1, R2, G, NTH>), dim3(idivup(nimg * ((h + 1) / 2), G)), dim3(NTH),
                       size_t(G) * (L + 1) * sizeof(cx<float>), st, img, nimg, h, w, tw_of<float>(e, L), mean, spec);
}

// `mean`: per-image mean subtracted on load (null for none)
void rowspec_fast(Engine &e, int L, const float *img, int nimg, int h, int w, const double *mean, cx<float> *spec,
                  hipStream_t st) {
    switch (L) {
        case 72: rowspec_launch<8, 9>(e, img, nimg, h, w, mean, spec, st); break;
        case 80: rowspec_launch<16, 5>(e, img, nimg, h, w, mean, spec, st); break;
        case 104: rowspec_launch<8, 13>(e, img, nimg, h, w, mean, spec, st); break;
        case 128: rowspec_launch<16, 8>(e, img, nimg, h, w, mean, spec, st); break;
        case 144: rowspec_launch<16, 9>(e, img, nimg, h, w, mean, spec, st); break;
        case 168: rowspec_launch<8, 21>(e, img, nimg, h, w, mean, spec, st); break;
        case 195: rowspec_launch<13, 15>(e, img, nimg, h, w, mean, spec, st); break;
        case 256: rowspec_launch<16, 16>(e, img, nimg, h, w, mean, spec, st); break;
        case 272: rowspec_launch<16, 17>(e, img, nimg, h, w, mean, spec, st); break;
        case 289: rowspec_launch<17, 17>(e, img, nimg, h, w, mean, spec, st); break;
        case 325: rowspec_launch<13, 25>(e, img, nimg, h, w, mean, spec, st); break;
        case 528: rowspec_launch<16, 33, 7, 256>(e, img, nimg, h, w, mean, spec, st); break;
        default: throw Error(AMPCOR_E_ARG, "no in-place row transform of length " + std::to_string(L));
    }
}

// per-window copies of a chunk's outputs (pairs of int32, one T -> float); a
// null destination or source skips that output
template <class T>
__global__ void k_chunk_out(int n, const int32_t *loc, int32_t *o_loc, const int32_t *start, int32_t *o_start,
                            const int32_t *shift, int32_t *o_shift, const int32_t *zidx, int32_t *o_zidx, const T *val,
                            float *o_val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto pair = [&](const int32_t *s, int32_t *d) {
        if (s && d) { d[2 * i] = s[2 * i]; d[2 * i + 1] = s[2 * i + 1]; }
    };
    pair(loc, o_loc);
    pair(start, o_start);
    pair(shift, o_shift);
    pair(zidx, o_zidx);
    if (val && o_val) o_val[i] = float(val[i]);
}

// start[i], shift[i] = the given values for windows f0 <= i < n
__global__ void k_given_starts(const int32_t *gs, const int32_t *gh, int f0, int n, int32_t *start, int32_t *shift) {
    const int i = f0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    start[2 * i] = gs[2 * i];
    start[2 * i + 1] = gs[2 * i + 1];
    shift[2 * i] = gh[2 * i];
    shift[2 * i + 1] = gh[2 * i + 1];
}

inline dim3 grid2(int per_img, int nimg) { return dim3(std::min(idivup(per_img, 256), 64), nimg); }
}  // namespace

// cuAmpcorChunk::run (cuAmpcorChunk.cpp:12-217) over chunks of windows.
template <class T>
void Engine::run_staged_t(const RunArgs &a) {
    const std::string tn = sizeof(T) == 8 ? "d_" : "f_";
    const hipStream_t st = a.stream;
    const int W_h = p.window_h, W_w = p.window_w;
    const int S_h = sz.search_h, S_w = sz.search_w;
    const int chunk = std::min(a.chunk > 0 ? a.chunk : 1024, 65535);
    const int mode = p.image_complex ? (p.deramp == 0 ? 1 : 0) : 2;
    // correlation transform sizes (5-smooth, >= the search chips)
    const int L1h = sz.fft_corr_raw_h, L1w = sz.fft_corr_raw_w, L2h = sz.fft_corr_os_h, L2w = sz.fft_corr_os_w;
    const int Z = sz.zoom, So = sz.surf_out;
    const int nb = std::min(chunk, a.n);

    cx<T> *c_ref = buf<cx<T>>(tn + "c_ref", size_t(nb) * W_h * W_w);
    T *r_ref = buf<T>(tn + "r_ref", size_t(nb) * W_h * W_w);
    cx<T> *c_sec = buf<cx<T>>(tn + "c_sec", size_t(nb) * S_h * S_w);
    T *r_sec = buf<T>(tn + "r_sec", size_t(nb) * S_h * S_w);
    auto norm_lds = [](int sw) { return size_t(4 * sw + 2) * sizeof(double); };
    // block size of the one-block-per-window reductions
    auto per_window_threads = [](int size) { return size >= 16384 ? 1024 : 256; };
    double *ssd_raw = buf<double>(tn + "ssd_raw", nb), *ssd_os = buf<double>(tn + "ssd_os", nb);
    const size_t vmax = std::max(size_t(sz.raw_corr_h) * S_w, size_t(sz.os_corr_h) * sz.os_search_w);
    double *V1 = buf<double>(tn + "V1", size_t(nb) * vmax), *V2 = buf<double>(tn + "V2", size_t(nb) * vmax);
    // NCC normalization: barrier-free column/row passes up to 1024 search columns, else one block per window
    auto normalize = [&](T *c, int ch, int cw, const T *tm, int th, int tw, const T *sr, int sh, int sw,
                         const double *ssd, int nw) {
        if (sw <= 1024) {
            hipLaunchKernelGGL(k_colsums<T>, dim3(idivup(sw, 64), nw), dim3(64), 0, st, sr, sh, sw, th, ch, V1, V2);
            const dim3 g(idivup(nw * ch, 4)), b(256);
            if (sw <= 128) hipLaunchKernelGGL((k_norm_rows<T, 2>), g, b, 0, st, c, nw, ch, cw, th, tw, V1, V2, sw, ssd);
            else if (sw <= 256) hipLaunchKernelGGL((k_norm_rows<T, 4>), g, b, 0, st, c, nw, ch, cw, th, tw, V1, V2, sw, ssd);
            else if (sw <= 576) hipLaunchKernelGGL((k_norm_rows<T, 9>), g, b, 0, st, c, nw, ch, cw, th, tw, V1, V2, sw, ssd);
            else hipLaunchKernelGGL((k_norm_rows<T, 16>), g, b, 0, st, c, nw, ch, cw, th, tw, V1, V2, sw, ssd);
        } else {
            hipLaunchKernelGGL(k_normalize<T>, dim3(nw), dim3(256), norm_lds(sw), st, c, ch, cw, tm, th, tw, sr, sh, sw);
        }
    };
    T *corr = buf<T>(tn + "corr", size_t(nb) * sz.raw_corr_h * sz.raw_corr_w);
    int32_t *loc = buf<int32_t>(tn + "loc", size_t(nb) * 2);
    T *val = buf<T>(tn + "val", nb);
    int32_t *start = buf<int32_t>(tn + "start", size_t(nb) * 2);
    int32_t *shift = buf<int32_t>(tn + "shift", size_t(nb) * 2);
    int32_t *zidx = buf<int32_t>(tn + "zidx", size_t(nb) * 2);
    cx<T> *c_ref_os = buf<cx<T>>(tn + "c_ref_os", size_t(nb) * sz.os_win_h * sz.os_win_w);
    T *r_ref_os = buf<T>(tn + "r_ref_os", size_t(nb) * sz.os_win_h * sz.os_win_w);
    cx<T> *c_zoom = buf<cx<T>>(tn + "c_zoom", size_t(nb) * sz.zoom_in_h * sz.zoom_in_w);
    // deramp rotation tables (row then column factors per chip)
    cx<double> *dtab = p.deramp == 1 ? buf<cx<double>>(tn + "dtab", size_t(nb) * (std::max(W_h + W_w,
                                                                 sz.zoom_in_h + sz.zoom_in_w))) : nullptr;
    const int OSh = sz.os_search_h, OSw = sz.os_search_w;   // oversampled zoom chip (fixed by the interpolation)
    // the separable oversampler covers the x2 interpolation of even-sided chips (an
    // odd side loses its middle spectrum line to the quadrant padding, which does
    // not factor into the half-sample form)
    const bool sep_os = oversampler >= 2 && p.raw_ovs == 2 && sz.os_win_h == 2 * W_h && sz.os_win_w == 2 * W_w &&
                        OSh == 2 * sz.zoom_in_h && OSw == 2 * sz.zoom_in_w && W_h % 2 == 0 && W_w % 2 == 0 &&
                        sz.zoom_in_h % 2 == 0 && sz.zoom_in_w % 2 == 0;
    cx<T> *c_sec_os = buf<cx<T>>(tn + "c_sec_os", size_t(nb) * OSh * OSw);
    T *r_sec_os = buf<T>(tn + "r_sec_os", size_t(nb) * OSh * OSw);
    T *corr_os = buf<T>(tn + "corr_os", size_t(nb) * sz.os_corr_h * sz.os_corr_w);
    // Correlations through row spectra (k_rowspec, k_mac_lags, k_lagrows_real): row
    // FFTs of the template and search rows, per-bin sums over template rows for each
    // lag row, and row IFFTs of the lag rows only. The padded 2-D FFT route remains
    // for shapes whose row pairs outnumber the transform rows.
    auto fits_rows = [&](int th, int sh, int Lh) { return (th + 1) / 2 + (sh + 1) / 2 <= Lh; };
    const bool row_raw = fits_rows(W_h, S_h, L1h), row_os = fits_rows(sz.os_win_h, OSh, L2h);
    // row transform lengths: an in-place length for the FP32 passes where one is listed
    const bool f32 = sizeof(T) == 4;
    const int R1w = f32 && row_len_fast(S_w) ? row_len_fast(S_w) : L1w;
    const int R2w = f32 && row_len_fast(OSw) ? row_len_fast(OSw) : L2w;
    // the oversampled template's mean from the row kernel's block sums, subtracted
    // as the row spectra load it (no k_sub_mean pass), where every piece applies:
    // FP32, separable in-place rows, whole-chip blocks, the in-place row spectra,
    // and the normalization that needs no mean-free template image (search <= 1024)
    const bool fuse_mean = f32 && sep_os && oversampler == 2 && os2_rows_has(W_w) && os2_rows_has(W_h) &&
                           sz.os_win_h % os2_rows_per_block<T>(W_w) == 0 && row_os && R2w != L2w && OSw <= 1024;
    double *os_part = fuse_mean ? buf<double>(tn + "os_part", size_t(nb) * 2 * (sz.os_win_h / os2_rows_per_block<T>(W_w)))
                                : nullptr;
    double *os_mean = fuse_mean ? buf<double>(tn + "os_mean", nb) : nullptr;
    size_t zmax = 0;
    if (!row_raw) zmax = size_t(L1h) * L1w;
    if (!row_os) zmax = std::max(zmax, size_t(L2h) * L2w);
    cx<T> *zc = zmax ? buf<cx<T>>(tn + "zc", size_t(nb) * zmax) : nullptr;
    size_t spec_elems = 0, q_elems = 0;
    if (row_raw) {
        spec_elems = std::max(spec_elems, size_t(W_h + S_h) * (R1w / 2 + 1));
        q_elems = std::max(q_elems, size_t(sz.raw_corr_h) * R1w);
    }
    if (row_os) {
        spec_elems = std::max(spec_elems, size_t(sz.os_win_h + OSh) * (R2w / 2 + 1));
        q_elems = std::max(q_elems, size_t(sz.os_corr_h) * R2w);
    }
    cx<T> *rspec = spec_elems ? buf<cx<T>>(tn + "rspec", size_t(nb) * spec_elems) : nullptr;
    cx<T> *rq = q_elems ? buf<cx<T>>(tn + "rq", size_t(nb) * q_elems) : nullptr;
    // tmean: the template's mean, subtracted as its rows load (FP32 in-place row path), or null
    auto row_correlate = [&](const T *tm, int th, int tw, const T *sr, int sh, int sw, int L, int NL, int NLw,
                             T *out, int n, const double *tmean) {
        const int H1 = L / 2 + 1;
        const FftPlan &pl = plan(L);
        const FftDesc &d = sizeof(T) == 4 ? pl.desc32 : pl.desc64;
        const int g = std::max(1, fft_elems<T>(*this) / L);
        const size_t lds = 2 * size_t(g) * L * sizeof(cx<T>);
        cx<T> *tspec = rspec, *sspec = rspec + size_t(n) * th * H1;
        if (sizeof(T) == 4 && row_len_fast(sw) == L) {
            rowspec_fast(*this, L, reinterpret_cast<const float *>(tm), n, th, tw, tmean,
                         reinterpret_cast<cx<float> *>(tspec), st);
            rowspec_fast(*this, L, reinterpret_cast<const float *>(sr), n, sh, sw, nullptr,
                         reinterpret_cast<cx<float> *>(sspec), st);
        } else {
            hipLaunchKernelGGL(k_rowspec<T>, dim3(idivup(n * ((th + 1) / 2), g)), dim3(256), lds, st, tm, n, th, tw, g,
                               d, tspec);
            hipLaunchKernelGGL(k_rowspec<T>, dim3(idivup(n * ((sh + 1) / 2), g)), dim3(256), lds, st, sr, n, sh, sw, g,
                               d, sspec);
        }
        // lags per thread: 17 once there is more than one group of 9 (each group
        // re-reads every spectrum; the sums themselves do not depend on the
        // grouping), unless that leaves too few workgroups to fill the GPU (the
        // FP64 re-run of a few hundred windows)
        if (NL > 9 && n * idivup(H1, 64) >= 2048)
            hipLaunchKernelGGL((k_mac_lags<T, 17, 2>), dim3(idivup(H1, 64), idivup(NL, 17), n), dim3(128), 0, st, tspec,
                               th, sspec, sh, NL, L, rq, T(1.0 / double(L)));
        else
            hipLaunchKernelGGL((k_mac_lags<T, 9, 4>), dim3(idivup(H1, 64), idivup(NL, 9), n), dim3(256), 0, st, tspec,
                               th, sspec, sh, NL, L, rq, T(1.0 / double(L)));
        hipLaunchKernelGGL(k_lagrows_real<T>, dim3(idivup(n * NL, g)), dim3(256), lds, st, rq, n * NL, g, d, NLw, out);
    };
    // the 16-sparse surface kernel needs no 256x256 buffers
    const bool fast_surface = !p.surf_method && Z == 16 && So == 256;
    cx<T> *c_surf_in = fast_surface ? nullptr : buf<cx<T>>(tn + "c_surf_in", size_t(nb) * Z * Z);
    cx<T> *c_surf = fast_surface || p.surf_method ? nullptr : buf<cx<T>>(tn + "c_surf", size_t(nb) * So * So);
    T *r_surf = fast_surface ? nullptr : buf<T>(tn + "r_surf", size_t(nb) * So * So);
    T *r_adj = buf<T>(tn + "r_adj", size_t(nb) * Z * Z);
    if (p.surf_method) ensure_sinc_table();

    // windows from a.fine_from on have their coarse zoom start and shift given
    // (a re-run of windows whose coarse peak is not in doubt): chunks of only
    // such windows skip the raw stage; in a mixed chunk they are overwritten
    // after it
    const bool given = a.fine_start != nullptr;
    for (int w0 = 0; w0 < a.n; w0 += nb) {
        const int n = std::min(nb, a.n - w0);
        const bool fine_only = given && w0 >= a.fine_from;
        // chip load + amplitude + mean (cuAmpcorChunk.cpp:18-39)
        hipLaunchKernelGGL(k_gather<T>, grid2(W_h * W_w, n), dim3(256), 0, st, a.ref, a.ref_pitch,
                           a.ref_d + w0, a.ref_a + w0, W_h, W_w, mode, c_ref, fine_only ? nullptr : r_ref);
        hipLaunchKernelGGL(k_gather<T>, grid2(S_h * S_w, n), dim3(256), 0, st, a.sec, a.sec_pitch,
                           a.sec_d + w0, a.sec_a + w0, S_h, S_w, mode, c_sec, fine_only ? nullptr : r_sec);
        if (fine_only) {
            HIP_OK(hipMemcpyAsync(start, a.fine_start + 2 * size_t(w0), sizeof(int32_t) * 2 * n,
                                  hipMemcpyDeviceToDevice, st));
            HIP_OK(hipMemcpyAsync(shift, a.fine_shift + 2 * size_t(w0), sizeof(int32_t) * 2 * n,
                                  hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL(k_sub_mean<T>, dim3(n), dim3(per_window_threads(W_h * W_w)), 0, st, r_ref, W_h * W_w,
                               ssd_raw);
            // raw correlation (cuAmpcorChunk.cpp:48-52)
            if (row_raw) {
                row_correlate(r_ref, W_h, W_w, r_sec, S_h, S_w, R1w, sz.raw_corr_h, sz.raw_corr_w, corr, n, nullptr);
            } else {
                hipLaunchKernelGGL(k_pack_corr<T>, grid2(L1h * L1w, n), dim3(256), 0, st, r_ref, W_h, W_w, r_sec,
                                   S_h, S_w, zc, L1h, L1w);
                fft2<T, -1>(*this, zc, n, L1h, L1w, st);
                hipLaunchKernelGGL(k_spec_product<T>, grid2(L1h * L1w, n), dim3(256), 0, st, zc, L1h, L1w,
                                   T(1.0 / (double(L1h) * L1w)));
                fft2_rows_kept<T, +1>(*this, zc, n, L1h, L1w, sz.raw_corr_h, st);
                hipLaunchKernelGGL(k_crop_real<T>, grid2(sz.raw_corr_h * sz.raw_corr_w, n), dim3(256), 0, st, zc,
                                   L1h, L1w, corr, sz.raw_corr_h, sz.raw_corr_w);
            }
            // normalize (cuAmpcorChunk.cpp:60)
            normalize(corr, sz.raw_corr_h, sz.raw_corr_w, r_ref, W_h, W_w, r_sec, S_h, S_w, ssd_raw, n);
            // argmax, covariance, SNR, zoom start/shift (cuAmpcorChunk.cpp:69-106)
            hipLaunchKernelGGL(k_maxloc<T>, dim3(n), dim3(256), 0, st, corr, sz.raw_corr_h, sz.raw_corr_w, loc, val,
                               a.margin_coarse ? a.margin_coarse + w0 : nullptr);
            hipLaunchKernelGGL(k_stats<T>, dim3(n), dim3(64), 0, st, corr, sz.raw_corr_h, sz.raw_corr_w, n,
                               loc, val, W_h * W_w, sz.corr_raw_zoom_h, sz.corr_raw_zoom_w, p.hs_down, p.hs_across,
                               sz.half_zoom, a.cov + 3 * size_t(w0), a.snr + w0, start, shift);
            if (given && a.fine_from < w0 + n) {
                const int f0 = a.fine_from - w0;
                hipLaunchKernelGGL(k_given_starts, dim3(idivup(n - f0, 256)), dim3(256), 0, st,
                                   a.fine_start + 2 * size_t(w0), a.fine_shift + 2 * size_t(w0), f0, n, start, shift);
            }
        }
        // reference oversampling (cuAmpcorChunk.cpp:115-126)
        if (p.deramp == 1)
            hipLaunchKernelGGL(k_deramp_phase<T>, dim3(n), dim3(per_window_threads(W_h * W_w)), 0, st, c_ref, W_h, W_w,
                               (const int32_t *)nullptr, W_h, W_w, dtab);
        if (sep_os) {
            os2_separable<T>(*this, c_ref, W_h, W_w, nullptr, p.deramp == 1 ? dtab : nullptr, n, W_h, W_w, c_ref_os,
                             r_ref_os, st, fuse_mean ? os_part : nullptr);
        } else {
            fft2_chip<T, +1>(*this, c_ref, W_h, W_w, nullptr, p.deramp == 1 ? dtab : nullptr, c_ref, n, W_h, W_w, st);
            fft2_quad<T, -1, 1>(*this, c_ref, n, W_h, W_w, c_ref_os, sz.os_win_h, sz.os_win_w, r_ref_os, st);
        }
        if (fuse_mean)
            hipLaunchKernelGGL(k_mean_parts, dim3(idivup(n, 256)), dim3(256), 0, st, os_part,
                               sz.os_win_h / os2_rows_per_block<T>(W_w), sz.os_win_h * sz.os_win_w, n, os_mean, ssd_os);
        else
            hipLaunchKernelGGL(k_sub_mean<T>, dim3(n), dim3(per_window_threads(sz.os_win_h * sz.os_win_w)), 0, st,
                               r_ref_os, sz.os_win_h * sz.os_win_w, ssd_os);
        // secondary zoom-in crop and oversampling (cuAmpcorChunk.cpp:134-137)
        if (p.deramp == 1)
            hipLaunchKernelGGL(k_deramp_phase<T>, dim3(n), dim3(per_window_threads(sz.zoom_in_h * sz.zoom_in_w)), 0,
                               st, c_sec, S_h, S_w, start, sz.zoom_in_h, sz.zoom_in_w, dtab);
        if (sep_os) {
            os2_separable<T>(*this, c_sec, S_h, S_w, start, p.deramp == 1 ? dtab : nullptr, n, sz.zoom_in_h,
                             sz.zoom_in_w, c_sec_os, r_sec_os, st);
        } else {
            fft2_chip<T, +1>(*this, c_sec, S_h, S_w, start, p.deramp == 1 ? dtab : nullptr, c_zoom, n, sz.zoom_in_h,
                             sz.zoom_in_w, st);
            fft2_quad<T, -1, 1>(*this, c_zoom, n, sz.zoom_in_h, sz.zoom_in_w, c_sec_os, OSh, OSw, r_sec_os, st);
        }
        // oversampled correlation + normalize (cuAmpcorChunk.cpp:148-161)
        if (row_os) {
            row_correlate(r_ref_os, sz.os_win_h, sz.os_win_w, r_sec_os, OSh, OSw, R2w, sz.os_corr_h, sz.os_corr_w,
                          corr_os, n, fuse_mean ? os_mean : nullptr);
        } else {
            hipLaunchKernelGGL(k_pack_corr<T>, grid2(L2h * L2w, n), dim3(256), 0, st, r_ref_os, sz.os_win_h,
                               sz.os_win_w, r_sec_os, OSh, OSw, zc, L2h, L2w);
            fft2<T, -1>(*this, zc, n, L2h, L2w, st);
            hipLaunchKernelGGL(k_spec_product<T>, grid2(L2h * L2w, n), dim3(256), 0, st, zc, L2h, L2w,
                               T(1.0 / (double(L2h) * L2w)));
            fft2_rows_kept<T, +1>(*this, zc, n, L2h, L2w, sz.os_corr_h, st);
            hipLaunchKernelGGL(k_crop_real<T>, grid2(sz.os_corr_h * sz.os_corr_w, n), dim3(256), 0, st, zc, L2h, L2w,
                               corr_os, sz.os_corr_h, sz.os_corr_w);
        }
        normalize(corr_os, sz.os_corr_h, sz.os_corr_w, r_ref_os, sz.os_win_h, sz.os_win_w, r_sec_os, OSh, OSw,
                  ssd_os, n);
        // trim + surface oversampling + argmax + sub-pixel offset (cuAmpcorChunk.cpp:169-207)
        float *mf = a.margin_fine ? a.margin_fine + w0 : nullptr;
        const float ratio = 1.0f / float(p.surf_ovs * p.raw_ovs);
        if (fast_surface) {
            // 16-sparse spectrum: one kernel instead of trim, FFT, pad, 256x256 FFT, real part, argmax
            const size_t lds = sizeof(cx<T>) * (2 * 256 + 16 * 256);
            hipLaunchKernelGGL(k_surface<T>, dim3(n), dim3(256), lds, st, corr_os, sz.os_corr_h, sz.os_corr_w,
                               start, p.hs_down, p.hs_across, ratio, static_cast<const cx<T> *>(plan(256).tw(T())),
                               zidx, mf, a.offset + 2 * size_t(w0));
        } else {
            if (p.surf_method) {
                hipLaunchKernelGGL(k_trim_real<T>, grid2(Z * Z, n), dim3(256), 0, st, corr_os, sz.os_corr_h,
                                   sz.os_corr_w, r_adj, Z);
                hipLaunchKernelGGL(k_sinc<T>, dim3(1, n), dim3(256), 0, st, r_adj, Z, r_surf, So, shift,
                                   p.surf_ovs * p.raw_ovs, sinc_table.as<float>(), p.surf_ovs, sinc_dec, sinc_len);
            } else {
                hipLaunchKernelGGL(k_trim_complex<T>, grid2(Z * Z, n), dim3(256), 0, st, corr_os, sz.os_corr_h,
                                   sz.os_corr_w, c_surf_in, Z);
                fft2<T, +1>(*this, c_surf_in, n, Z, Z, st);
                fft2_quad<T, -1, 2>(*this, c_surf_in, n, Z, Z, c_surf, So, So, r_surf, st);
            }
            hipLaunchKernelGGL(k_maxloc<T>, dim3(n), dim3(256), 0, st, r_surf, So, So, zidx, (T *)nullptr, mf);
            hipLaunchKernelGGL(k_subpixel, dim3(idivup(n, 64)), dim3(64), 0, st, start, zidx, n, ratio, p.hs_down,
                               p.hs_across, a.offset + 2 * size_t(w0));
        }
        if (a.refined)
            hipLaunchKernelGGL(k_refine<T>, dim3(n), dim3(64), refine_lds(Z), st, corr_os, sz.os_corr_h, sz.os_corr_w,
                               Z, p.surf_ovs, 1.0 / double(p.surf_ovs * p.raw_ovs), zidx, start, p.hs_down,
                               p.hs_across, static_cast<const cx<double> *>(plan(So).tw64.ptr),
                               a.refined + 2 * size_t(w0));
        // the chunk's integer results and peak value to the caller's arrays, one launch
        {
            auto at = [&](int32_t *o) { return o ? o + 2 * size_t(w0) : nullptr; };
            hipLaunchKernelGGL(k_chunk_out<T>, dim3(idivup(n, 256)), dim3(256), 0, st, n, fine_only ? nullptr : loc,
                               at(a.maxloc), start, at(a.start), shift, at(a.shift), zidx, at(a.zoom_idx),
                               fine_only ? nullptr : val, a.maxval ? a.maxval + w0 : nullptr);
        }
        HIP_OK(hipGetLastError());
    }
}

// FP32 staged pipeline; under precision 1 the windows whose argmax margins
// fall below the thresholds are re-run in FP64
void Engine::run_staged(const RunArgs &a) {
    RunArgs b = a;
    if (precision == 1) {
        if (!b.margin_coarse) b.margin_coarse = buf<float>("f_mc", a.n);
        if (!b.margin_fine) b.margin_fine = buf<float>("f_mf", a.n);
    }
    run_staged_t<float>(b);
    if (precision == 1 && !a.nosync) rerun_fp64(b, b.margin_coarse, b.margin_fine);
}
template void Engine::run_staged_t<double>(const RunArgs &a);

namespace {
// the four start arrays of the m selected windows into st4[0..4m), and, when
// given, their zoom starts and shifts (pairs) into st4[4m..6m) and st4[6m..8m)
__global__ void k_take_starts(const int32_t *rd, const int32_t *ra, const int32_t *sd, const int32_t *sa,
                              const int32_t *start32, const int32_t *shift32, const int32_t *idx, int m,
                              int32_t *st4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int w = idx[i];
    st4[i] = rd[w];
    st4[m + i] = ra[w];
    st4[2 * m + i] = sd[w];
    st4[3 * m + i] = sa[w];
    if (start32) {
        st4[4 * m + 2 * i] = start32[2 * w];
        st4[4 * m + 2 * i + 1] = start32[2 * w + 1];
        st4[6 * m + 2 * i] = shift32[2 * w];
        st4[6 * m + 2 * i + 1] = shift32[2 * w + 1];
    }
}

// Windows to re-run in FP64, in order: first every window whose coarse margin
// is below tc (or whose fine margin is below tf when no zoom starts are given:
// has_start 0), then every other window whose fine margin is below tf (NaN
// counts as below). cnt[0] = first kind, cnt[1] = both kinds. Two launches over
// blocks of SEL_WIN windows (four consecutive ones per thread): the per-block
// counts of both kinds, then each block's slots from the counts of the blocks
// before it and a block prefix sum of the per-thread counts.
constexpr int SEL_NT = 256, SEL_WIN = 4 * SEL_NT;

__device__ __forceinline__ void rerun_flags(const float *mc, const float *mf, int n, float tc, float tf,
                                            int has_start, int i0, unsigned &full, unsigned &fine) {
    full = fine = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = i0 + j;
        if (i < n) {
            const bool c = !(mc[i] >= tc), f = !(mf[i] >= tf);
            const bool w = c || (f && !has_start);
            if (w) full |= 1u << j;
            else if (f) fine |= 1u << j;
        }
    }
}

// sum of v over the workgroup (SEL_NT threads), in every thread
__device__ __forceinline__ int sel_block_sum(int v, int *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < SEL_NT / 64; ++w) t += sh[w];
    return t;
}

__global__ void __launch_bounds__(SEL_NT) k_select_count(const float *mc, const float *mf, int n, float tc, float tf,
                                                         int has_start, int32_t *bcnt) {
    __shared__ int sh[SEL_NT / 64];
    unsigned full, fine;
    rerun_flags(mc, mf, n, tc, tf, has_start, blockIdx.x * SEL_WIN + 4 * threadIdx.x, full, fine);
    const int cf = sel_block_sum(__popc(full), sh), cq = sel_block_sum(__popc(fine), sh);
    if (threadIdx.x == 0) {
        bcnt[2 * blockIdx.x] = cf;
        bcnt[2 * blockIdx.x + 1] = cq;
    }
}

__global__ void __launch_bounds__(SEL_NT) k_select_write(const float *mc, const float *mf, int n, float tc, float tf,
                                                         int has_start, const int32_t *bcnt, int nb, int32_t *idx,
                                                         int32_t *cnt) {
    __shared__ int sh[SEL_NT / 64], wf[SEL_NT / 64], wq[SEL_NT / 64];
    // slots of this block: first kind after the earlier blocks' first kind, second
    // kind after every first-kind window and the earlier blocks' second kind
    int pf = 0, pq = 0, tf_ = 0, tq = 0;
    for (int j = threadIdx.x; j < nb; j += SEL_NT) {
        const int a = bcnt[2 * j], b = bcnt[2 * j + 1];
        tf_ += a; tq += b;
        if (j < int(blockIdx.x)) { pf += a; pq += b; }
    }
    pf = sel_block_sum(pf, sh); pq = sel_block_sum(pq, sh);
    tf_ = sel_block_sum(tf_, sh); tq = sel_block_sum(tq, sh);
    const int i0 = blockIdx.x * SEL_WIN + 4 * threadIdx.x;
    unsigned full, fine;
    rerun_flags(mc, mf, n, tc, tf, has_start, i0, full, fine);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int mf_ = __popc(full), mq = __popc(fine);
    int xf = mf_, xq = mq;                 // inclusive prefix sums over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int a = __shfl_up(xf, o), b = __shfl_up(xq, o);
        if (lane >= o) { xf += a; xq += b; }
    }
    if (lane == 63) { wf[wv] = xf; wq[wv] = xq; }
    __syncthreads();
    int of = 0, oq = 0;
#pragma unroll
    for (int w = 0; w < SEL_NT / 64; ++w)
        if (w < wv) { of += wf[w]; oq += wq[w]; }
    int sf = pf + of + xf - mf_, sq = tf_ + pq + oq + xq - mq;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (full & (1u << j)) idx[sf++] = i0 + j;
        if (fine & (1u << j)) idx[sq++] = i0 + j;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[0] = tf_;
        cnt[1] = tf_ + tq;
    }
}

// the first window (lowest index) whose chips leave their images, or n, into *first
// (initialized to n by the caller's launch of k_set_i32)
__global__ void k_first_bad(const int32_t *rd, const int32_t *ra, const int32_t *sd, const int32_t *sa, int n,
                            int64_t ref_h, int64_t ref_w, int64_t sec_h, int64_t sec_w, int wh, int ww, int sh,
                            int sw, int32_t *first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = rd[i], c = ra[i], s = sd[i], t = sa[i];
    if (r < 0 || c < 0 || r + wh > ref_h || c + ww > ref_w || s < 0 || t < 0 || s + sh > sec_h || t + sw > sec_w)
        atomicMin(first, i);
}

__global__ void k_set_i32(int32_t *p, int v) { *p = v; }
template <class U>
__global__ void k_put(const U *src, const int32_t *idx, U *dst, int n, int width) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * width) dst[size_t(idx[i / width]) * width + i % width] = src[i];
}
}  // namespace

// Re-run in FP64 (staged pipeline, T = double) every window whose coarse or
// fine argmax margin (top-1 minus top-2 of the normalized surface) is below
// its threshold, and overwrite that window's outputs. FP32 rounding can only
// reorder near-equal candidates, so this makes the integer peaks those of an
// FP64 evaluation. When the coarse zoom start and shift of every window are
// given (start32, shift32) and only a window's fine margin is small, its fine
// stage is re-run from that start and its coarse peak, SNR and covariance stand
// as computed. Both kinds go through one staged FP64 pass (one launch per
// stage for all of them: the pass is latency-bound at these window counts).
void Engine::rerun_fp64(const RunArgs &a, const float *mc, const float *mf, const int32_t *start32,
                        const int32_t *shift32) {
    // the windows to re-run, selected on the device: those needing the whole
    // pipeline first, then those re-run from the zoom stage on (their coarse
    // results, SNR and covariance stay the FP32 ones); only the two counts
    // come back to the host
    // p_count[2]: the first window whose chips leave their images (a.n if none),
    // set by Engine::run's bounds check
    int32_t *d_idx = buf<int32_t>("p_idx", a.n);
    int32_t *d_cnt = buf<int32_t>("p_count", 3);
    const int nb = idivup(a.n, SEL_WIN);
    int32_t *d_bcnt = buf<int32_t>("p_bcount", 2 * size_t(nb));
    hipLaunchKernelGGL(k_select_count, dim3(nb), dim3(SEL_NT), 0, a.stream, mc, mf, a.n, tau_coarse, tau_fine,
                       start32 != nullptr, d_bcnt);
    hipLaunchKernelGGL(k_select_write, dim3(nb), dim3(SEL_NT), 0, a.stream, mc, mf, a.n, tau_coarse, tau_fine,
                       start32 != nullptr, d_bcnt, nb, d_idx, d_cnt);
    int32_t cnt[3];
    HIP_OK(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, a.stream));
    HIP_OK(hipStreamSynchronize(a.stream));
    if (cnt[2] < a.n) throw Error(AMPCOR_E_RANGE, "window " + std::to_string(cnt[2]) + " reads outside its image");
    const int m_full = cnt[0], m = cnt[1];
    prec_checked += a.n;
    prec_rerun += m;
    if (m == 0) return;
    // one FP64 pass over both kinds
    int32_t *st4 = buf<int32_t>("p_starts", size_t(8) * m);
    hipLaunchKernelGGL(k_take_starts, dim3(idivup(m, 256)), dim3(256), 0, a.stream, a.ref_d, a.ref_a, a.sec_d,
                       a.sec_a, m_full < m ? start32 : nullptr, m_full < m ? shift32 : nullptr, d_idx, m, st4);
    float *o_off = buf<float>("p_off", size_t(2) * m), *o_snr = buf<float>("p_snr", m);
    float *o_cov = buf<float>("p_cov", size_t(3) * m), *o_val = buf<float>("p_val", m);
    int32_t *o_i = buf<int32_t>("p_ints", size_t(8) * m);
    RunArgs sub = a;
    sub.ref_d = st4; sub.ref_a = st4 + m; sub.sec_d = st4 + 2 * m; sub.sec_a = st4 + 3 * m;
    sub.n = m;
    sub.margin_coarse = sub.margin_fine = nullptr;
    sub.offset = o_off; sub.snr = o_snr; sub.cov = o_cov; sub.maxval = o_val;
    sub.maxloc = o_i; sub.start = o_i + 2 * m; sub.shift = o_i + 4 * m; sub.zoom_idx = o_i + 6 * m;
    sub.chunk = 0;
    sub.fine_start = sub.fine_shift = nullptr;
    double *o_ref = a.refined ? buf<double>("p_refined", size_t(2) * m) : nullptr;
    sub.refined = o_ref;
    if (m_full < m) {
        sub.fine_start = st4 + 4 * m;
        sub.fine_shift = st4 + 6 * m;
        sub.fine_from = m_full;
    }
    run_staged_t<double>(sub);
    // offsets and zoom peaks of every re-run window; the rest of the full ones only
    auto put_f = [&](const float *src, float *dst, int cnt, int width) {
        if (dst && cnt) hipLaunchKernelGGL(k_put<float>, dim3(idivup(cnt * width, 256)), dim3(256), 0, a.stream, src, d_idx, dst, cnt, width);
    };
    auto put_i = [&](const int32_t *src, int32_t *dst, int cnt) {
        if (dst && cnt) hipLaunchKernelGGL(k_put<int32_t>, dim3(idivup(cnt * 2, 256)), dim3(256), 0, a.stream, src, d_idx, dst, cnt, 2);
    };
    put_f(o_off, a.offset, m, 2);
    if (o_ref)
        hipLaunchKernelGGL(k_put<double>, dim3(idivup(m * 2, 256)), dim3(256), 0, a.stream, o_ref, d_idx, a.refined, m, 2);
    put_i(sub.zoom_idx, a.zoom_idx, m);
    put_f(o_snr, a.snr, m_full, 1);
    put_f(o_cov, a.cov, m_full, 3);
    put_f(o_val, a.maxval, m_full, 1);
    put_i(sub.maxloc, a.maxloc, m_full);
    put_i(sub.start, a.start, m_full);
    put_i(sub.shift, a.shift, m_full);
    HIP_OK(hipGetLastError());
}

void Engine::run(const RunArgs &in) {
    if (!configured) throw Error(AMPCOR_E_ARG, "ampcor_setup has not been called");
    if (in.n < 0) throw Error(AMPCOR_E_ARG, "negative window count");
    if (in.n == 0) return;
    if (!in.ref || !in.sec || !in.ref_d || !in.ref_a || !in.sec_d || !in.sec_a || !in.offset || !in.snr || !in.cov)
        throw Error(AMPCOR_E_ARG, "null pointer argument");
    if (in.ref_pitch < in.ref_w || in.sec_pitch < in.sec_w) throw Error(AMPCOR_E_ARG, "pitch smaller than width");
    if ((in.mode & ~(0xff | AMPCOR_FLAG_NOSYNC)) != 0 || (in.mode & 0xff) > AMPCOR_MODE_FUSED)
        throw Error(AMPCOR_E_ARG, "unknown mode");
    HIP_OK(hipSetDevice(device));
    RunArgs a = in;
    a.mode = in.mode & 0xff;
    a.nosync = (in.mode & AMPCOR_FLAG_NOSYNC) != 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_OK(hipStreamIsCapturing(a.stream, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (capturing && !a.nosync) throw Error(AMPCOR_E_ARG, "a call under stream capture needs AMPCOR_FLAG_NOSYNC");
    if (capturing && profiling) throw Error(AMPCOR_E_ARG, "disable ampcor_profile before capturing");
    struct Restore { bool &f; bool v; ~Restore() { f = v; } } restore{no_alloc, no_alloc};
    no_alloc = capturing;
    if (capturing) captured = true;

    if (a.refined && (p.surf_method != 0 || sz.zoom > 64))
        throw Error(AMPCOR_E_UNSUPPORTED, "the FP64 peak refinement needs FFT surface oversampling and a zoom "
                                          "window of at most 64");
    const bool fused = (a.mode == AMPCOR_MODE_FUSED) || (a.mode == AMPCOR_MODE_AUTO && sz.fused);
    if (a.mode == AMPCOR_MODE_FUSED && !sz.fused)
        throw Error(AMPCOR_E_UNSUPPORTED, "the fused kernels do not cover this configuration");
    if (a.ref_h < p.window_h || a.ref_w < p.window_w || a.sec_h < sz.search_h || a.sec_w < sz.search_w)
        throw Error(AMPCOR_E_RANGE, "image smaller than one chip");
    int32_t *bad = nullptr;
    if (a.nosync) {
        // device-side bounds check: out-of-image windows read window (0, 0)
        // instead and report NaN outputs
        int32_t *st4 = buf<int32_t>("ns_starts", 4 * size_t(a.n));
        bad = buf<int32_t>("ns_bad", a.n);
        hipLaunchKernelGGL(k_sanitize_starts<int32_t>, dim3(idivup(a.n, 256)), dim3(256), 0, a.stream, a.ref_d, a.ref_a,
                           a.sec_d, a.sec_a, a.n, a.ref_h, a.ref_w, a.sec_h, a.sec_w, p.window_h, p.window_w,
                           sz.search_h, sz.search_w, st4, bad);
        a.ref_d = st4; a.ref_a = st4 + a.n; a.sec_d = st4 + 2 * size_t(a.n); a.sec_a = st4 + 3 * size_t(a.n);
    } else if (precision == 1) {
        // bounds checked on the device and reported with the FP64 re-run's counts
        // (one host synchronization per call): out-of-image windows read window
        // (0, 0) meanwhile, and the call fails with AMPCOR_E_RANGE before the re-run
        int32_t *cnt = buf<int32_t>("p_count", 3);
        int32_t *st4 = buf<int32_t>("ns_starts", 4 * size_t(a.n));
        int32_t *flags = buf<int32_t>("ns_bad", a.n);
        hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, a.stream, cnt + 2, int(a.n));
        hipLaunchKernelGGL(k_first_bad, dim3(idivup(a.n, 256)), dim3(256), 0, a.stream, a.ref_d, a.ref_a, a.sec_d,
                           a.sec_a, int(a.n), a.ref_h, a.ref_w, a.sec_h, a.sec_w, p.window_h, p.window_w,
                           sz.search_h, sz.search_w, cnt + 2);
        hipLaunchKernelGGL(k_sanitize_starts<int32_t>, dim3(idivup(a.n, 256)), dim3(256), 0, a.stream, a.ref_d, a.ref_a,
                           a.sec_d, a.sec_a, a.n, a.ref_h, a.ref_w, a.sec_h, a.sec_w, p.window_h, p.window_w,
                           sz.search_h, sz.search_w, st4, flags);
        a.ref_d = st4; a.ref_a = st4 + a.n; a.sec_d = st4 + 2 * size_t(a.n); a.sec_a = st4 + 3 * size_t(a.n);
    } else {
        // bounds: every chip must lie inside its image, checked on the device (the
        // first failing window comes back in one 4-byte copy; the caller's
        // checkPixelInImageRange is stricter)
        int32_t *first = buf<int32_t>("range_first", 1);
        hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, a.stream, first, int(a.n));
        hipLaunchKernelGGL(k_first_bad, dim3(idivup(a.n, 256)), dim3(256), 0, a.stream, a.ref_d, a.ref_a, a.sec_d,
                           a.sec_a, int(a.n), a.ref_h, a.ref_w, a.sec_h, a.sec_w, p.window_h, p.window_w,
                           sz.search_h, sz.search_w, first);
        int32_t bad_i = 0;
        HIP_OK(hipMemcpyAsync(&bad_i, first, sizeof(int32_t), hipMemcpyDeviceToHost, a.stream));
        HIP_OK(hipStreamSynchronize(a.stream));
        if (bad_i < a.n) throw Error(AMPCOR_E_RANGE, "window " + std::to_string(bad_i) + " reads outside its image");
    }
    if (precision == 2) run_staged_t<double>(a);
    else if (fused) run_fused(a);
    else run_staged(a);
    if (bad)
        hipLaunchKernelGGL(k_mark_bad<float>, dim3(idivup(a.n, 256)), dim3(256), 0, a.stream, bad, a.n, a.offset, a.snr, a.cov);
    HIP_OK(hipGetLastError());
}

// ampcor_run_windows_noalloc: the fused kernels over the device's static scratch
// array, FP32 only, bounds checked on the device, no host synchronization; any
// step that would allocate (a missing table) throws instead.
void Engine::run_noalloc(const RunArgs &a) {
    if (!configured) throw Error(AMPCOR_E_ARG, "ampcor_setup has not been called");
    if (a.n < 0) throw Error(AMPCOR_E_ARG, "negative window count");
    if (a.n == 0) return;
    if (!a.ref || !a.sec || !a.ref_d || !a.ref_a || !a.sec_d || !a.sec_a || !a.offset || !a.snr || !a.cov)
        throw Error(AMPCOR_E_ARG, "null pointer argument");
    if (a.ref_pitch < a.ref_w || a.sec_pitch < a.sec_w) throw Error(AMPCOR_E_ARG, "pitch smaller than width");
    if (!sz.fused)
        throw Error(AMPCOR_E_UNSUPPORTED, "the allocation-free entry runs the fused kernels only, which do not "
                                          "cover this configuration");
    if (a.ref_h < p.window_h || a.ref_w < p.window_w || a.sec_h < sz.search_h || a.sec_w < sz.search_w)
        throw Error(AMPCOR_E_RANGE, "image smaller than one chip");
    HIP_OK(hipSetDevice(device));
    struct Restore {
        bool &f, &g;
        bool v, w;
        ~Restore() { f = v; g = w; }
    } restore{no_alloc, profiling, no_alloc, profiling};
    no_alloc = true;
    profiling = false;
    run_fused_noalloc(a);
}

}  // namespace ampcor

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct ampcor_ctx {
    ampcor::Engine eng;
};

namespace {
std::mutex g_err_mu;
std::string g_err;
template <class F>
int guarded(ampcor_ctx *ctx, F &&f) {
    try {
        f();
        return AMPCOR_OK;
    } catch (const ampcor::Error &e) {
        if (ctx) ctx->eng.last_error = e.what();
        else { std::lock_guard<std::mutex> l(g_err_mu); g_err = e.what(); }
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->eng.last_error = e.what();
        else { std::lock_guard<std::mutex> l(g_err_mu); g_err = e.what(); }
        return AMPCOR_E_DEVICE;
    }
}
}  // namespace

extern "C" {

int ampcor_create(int32_t device, ampcor_ctx **out) {
    if (!out) return AMPCOR_E_ARG;
    *out = nullptr;
    return guarded(nullptr, [&] {
        int count = 0;
        HIP_OK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw ampcor::Error(AMPCOR_E_ARG, "device " + std::to_string(device) + " not available (" +
                                                  std::to_string(count) + " devices)");
        HIP_OK(hipSetDevice(device));
        // the allocation-free entry's scratch address on this device (once per device)
        void *scratch = ampcor::noalloc_scratch_init(device);
        auto *c = new ampcor_ctx();
        c->eng.device = device;
        c->eng.scratch_ptr = scratch;
        *out = c;
    });
}

int ampcor_select_near_ties(ampcor_ctx *ctx, const float *margin_coarse, const float *margin_fine, int32_t n,
                            float tau_coarse, float tau_fine, int32_t cap, int32_t *idx, void *stream) {
    if (!ctx) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        if (n < 0 || cap < 0 || cap > n) throw ampcor::Error(AMPCOR_E_ARG, "the capacity must lie in [0, n]");
        if (cap == 0) return;
        if (!margin_coarse || !margin_fine || !idx) throw ampcor::Error(AMPCOR_E_ARG, "null pointer argument");
        HIP_OK(hipSetDevice(ctx->eng.device));
        hipLaunchKernelGGL(ampcor::k_select_near, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream),
                           margin_coarse, margin_fine, n, tau_coarse, tau_fine, cap, idx);
        HIP_OK(hipGetLastError());
    });
}

int ampcor_run_windows_noalloc(ampcor_ctx *ctx, const void *ref_img, int64_t ref_h, int64_t ref_w,
                               int64_t ref_pitch, const void *sec_img, int64_t sec_h, int64_t sec_w,
                               int64_t sec_pitch, const int32_t *ref_d, const int32_t *ref_a, const int32_t *sec_d,
                               const int32_t *sec_a, int32_t n, float *out_offset, float *out_snr, float *out_cov,
                               void *stream) {
    if (!ctx) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        ampcor::Engine::RunArgs a{ref_img, ref_h, ref_w, ref_pitch, sec_img, sec_h, sec_w, sec_pitch,
                                  ref_d, ref_a, sec_d, sec_a, n, out_offset, out_snr, out_cov,
                                  nullptr, nullptr, nullptr, nullptr, nullptr,
                                  0, AMPCOR_MODE_FUSED, static_cast<hipStream_t>(stream)};
        a.nosync = true;
        ctx->eng.run_noalloc(a);
    });
}

int ampcor_destroy(ampcor_ctx *ctx) {
    if (!ctx) return AMPCOR_E_ARG;
    (void)hipSetDevice(ctx->eng.device);
    delete ctx;
    return AMPCOR_OK;
}

int ampcor_setup(ampcor_ctx *ctx, const ampcor_params *params, ampcor_sizes *sizes_out) {
    if (!ctx || !params) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        HIP_OK(hipSetDevice(ctx->eng.device));
        ctx->eng.setup(*params);
        if (sizes_out) *sizes_out = ctx->eng.sz;
    });
}

int ampcor_run_windows_debug(ampcor_ctx *ctx, const void *ref_img, int64_t ref_h, int64_t ref_w, int64_t ref_pitch,
                             const void *sec_img, int64_t sec_h, int64_t sec_w, int64_t sec_pitch,
                             const int32_t *ref_d, const int32_t *ref_a, const int32_t *sec_d,
                             const int32_t *sec_a, int32_t n, float *out_offset, float *out_snr, float *out_cov,
                             int32_t *out_maxloc, int32_t *out_start, int32_t *out_shift, int32_t *out_zoom_idx,
                             float *out_maxval, float *out_margin_coarse, float *out_margin_fine,
                             double *out_offset_fp64, int32_t chunk, int32_t mode, void *stream) {
    if (!ctx) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        ampcor::Engine::RunArgs a{ref_img, ref_h, ref_w, ref_pitch, sec_img, sec_h, sec_w, sec_pitch,
                                  ref_d, ref_a, sec_d, sec_a, n, out_offset, out_snr, out_cov,
                                  out_maxloc, out_start, out_shift, out_zoom_idx, out_maxval,
                                  chunk, mode, static_cast<hipStream_t>(stream),
                                  out_margin_coarse, out_margin_fine};
        a.refined = out_offset_fp64;
        ctx->eng.run(a);
    });
}

int ampcor_set_precision(ampcor_ctx *ctx, int32_t mode, float tau_coarse, float tau_fine) {
    if (!ctx || mode < 0 || mode > 2) return AMPCOR_E_ARG;
    ctx->eng.precision = mode;
    if (tau_coarse >= 0.f) ctx->eng.tau_coarse = tau_coarse;
    if (tau_fine >= 0.f) ctx->eng.tau_fine = tau_fine;
    return AMPCOR_OK;
}

int ampcor_set_option(ampcor_ctx *ctx, const char *name, int64_t value) {
    if (!ctx || !name) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        const std::string k(name);
        if (k == "fft_block_elems") {
            if (value != 1024 && value != 2048 && value != 4096 && value != 8192)
                throw ampcor::Error(AMPCOR_E_ARG, "fft_block_elems must be 1024, 2048, 4096 or 8192");
            ctx->eng.fft_block_elems = int(value);
        } else if (k == "fft_cols_width") {
            if (value != 0 && value != 8 && value != 16)
                throw ampcor::Error(AMPCOR_E_ARG, "fft_cols_width must be 0, 8 or 16");
            ctx->eng.fft_cols_width = int(value);
        } else if (k == "fine_kernel") {
            if (value != 1 && value != 2) throw ampcor::Error(AMPCOR_E_ARG, "fine_kernel must be 1 or 2");
            ctx->eng.fine_kernel = int(value);
        } else if (k == "oversampler") {
            if (value < 1 || value > 3) throw ampcor::Error(AMPCOR_E_ARG, "oversampler must be 1, 2 or 3");
            ctx->eng.oversampler = int(value);
        } else if (k == "fine_threads") {
            if (value != 512 && value != 1024) throw ampcor::Error(AMPCOR_E_ARG, "fine_threads must be 512 or 1024");
            ctx->eng.fine_threads = int(value);
        } else {
            throw ampcor::Error(AMPCOR_E_ARG, "unknown option " + k);
        }
    });
}

int ampcor_set_fft_radices(ampcor_ctx *ctx, int32_t n, const int32_t *radices, int32_t count) {
    if (!ctx) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        HIP_OK(hipSetDevice(ctx->eng.device));
        ctx->eng.set_fft_radices(n, radices, count);
    });
}

int ampcor_precision_stats(ampcor_ctx *ctx, int64_t *checked, int64_t *rerun) {
    if (!ctx) return AMPCOR_E_ARG;
    if (checked) *checked = ctx->eng.prec_checked;
    if (rerun) *rerun = ctx->eng.prec_rerun;
    ctx->eng.prec_checked = ctx->eng.prec_rerun = 0;
    return AMPCOR_OK;
}

int ampcor_run_windows(ampcor_ctx *ctx, const void *ref_img, int64_t ref_h, int64_t ref_w, int64_t ref_pitch,
                       const void *sec_img, int64_t sec_h, int64_t sec_w, int64_t sec_pitch, const int32_t *ref_d,
                       const int32_t *ref_a, const int32_t *sec_d, const int32_t *sec_a, int32_t n,
                       float *out_offset, float *out_snr, float *out_cov, int32_t chunk, int32_t mode,
                       void *stream) {
    return ampcor_run_windows_debug(ctx, ref_img, ref_h, ref_w, ref_pitch, sec_img, sec_h, sec_w, sec_pitch, ref_d,
                                    ref_a, sec_d, sec_a, n, out_offset, out_snr, out_cov, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, chunk, mode, stream);
}

int ampcor_profile(ampcor_ctx *ctx, int32_t enable) {
    if (!ctx) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        ctx->eng.prof_collect();
        ctx->eng.prof_acc.clear();
        ctx->eng.profiling = enable != 0;
    });
}

int ampcor_profile_only(ampcor_ctx *ctx, const char *name) {
    if (!ctx) return AMPCOR_E_ARG;
    return guarded(ctx, [&] { ctx->eng.prof_only = name ? name : ""; });
}

int ampcor_profile_report(ampcor_ctx *ctx, char *buf, int64_t len) {
    if (!ctx || !buf || len <= 0) return AMPCOR_E_ARG;
    return guarded(ctx, [&] {
        ctx->eng.prof_collect();
        std::string out;
        for (auto &kv : ctx->eng.prof_acc)
            out += kv.first + " " + std::to_string(kv.second.launches) + " " + std::to_string(kv.second.ms) + " " +
                   std::to_string(kv.second.windows) + "\n";
        const size_t n = std::min<size_t>(out.size(), size_t(len - 1));
        std::memcpy(buf, out.data(), n);
        buf[n] = 0;
    });
}

const char *ampcor_last_error(const ampcor_ctx *ctx) {
    if (ctx) return ctx->eng.last_error.c_str();
    std::lock_guard<std::mutex> l(g_err_mu);
    return g_err.c_str();
}

const char *ampcor_version(void) { return "isce2_amd 0.1.0 (gfx950)"; }

// cuAmpcorParameter::setStartPixels, grid overloads (cuAmpcorParameter.cpp:167-205):
// window i = row * n_across + col; each search axis uses its own half range.
int ampcor_start_pixels(const ampcor_params *params, int32_t n_down, int32_t n_across, int32_t start_down0,
                        int32_t start_across0, int32_t skip_down, int32_t skip_across, int32_t gross_down0,
                        int32_t gross_across0, const int32_t *gross_down, const int32_t *gross_across,
                        int32_t *ref_d, int32_t *ref_a, int32_t *sec_d, int32_t *sec_a) {
    return guarded(nullptr, [&] {
        if (!params || !ref_d || !ref_a || !sec_d || !sec_a)
            throw ampcor::Error(AMPCOR_E_ARG, "null pointer argument");
        if (n_down <= 0 || n_across <= 0 || int64_t(n_down) * n_across > INT32_MAX)
            throw ampcor::Error(AMPCOR_E_ARG, "window grid must be positive and below 2^31 windows");
        if ((gross_down == nullptr) != (gross_across == nullptr))
            throw ampcor::Error(AMPCOR_E_ARG, "give both varying gross offset arrays or neither");
        for (int64_t row = 0; row < n_down; ++row)
            for (int64_t col = 0; col < n_across; ++col) {
                const int64_t i = row * n_across + col;
                const int64_t rd = start_down0 + row * skip_down, ra = start_across0 + col * skip_across;
                const int64_t gd = gross_down ? gross_down[i] : gross_down0;
                const int64_t ga = gross_across ? gross_across[i] : gross_across0;
                const int64_t sd = rd + gd - params->hs_down, sa = ra + ga - params->hs_across;
                for (int64_t v : {rd, ra, sd, sa})
                    if (v < INT32_MIN || v > INT32_MAX)
                        throw ampcor::Error(AMPCOR_E_ARG, "start pixel of window " + std::to_string(i) +
                                                              " overflows int32");
                ref_d[i] = int32_t(rd); ref_a[i] = int32_t(ra);
                sec_d[i] = int32_t(sd); sec_a[i] = int32_t(sa);
            }
    });
}

// cuAmpcorParameter::checkPixelInImageRange (cuAmpcorParameter.cpp:275-331), the
// reference's strict rule: start >= 0 and start + chip size < image size.
int ampcor_check_range(const ampcor_params *params, const int32_t *ref_d, const int32_t *ref_a,
                       const int32_t *sec_d, const int32_t *sec_a, int32_t n, int64_t ref_h, int64_t ref_w,
                       int64_t sec_h, int64_t sec_w, int32_t *bad_window, int32_t *which) {
    if (bad_window) *bad_window = -1;
    if (which) *which = -1;
    return guarded(nullptr, [&] {
        if (!params || (n > 0 && (!ref_d || !ref_a || !sec_d || !sec_a)))
            throw ampcor::Error(AMPCOR_E_ARG, "null pointer argument");
        if (n < 0) throw ampcor::Error(AMPCOR_E_ARG, "negative window count");
        const int64_t wh = params->window_h, ww = params->window_w;
        const int64_t sh = wh + 2 * int64_t(params->hs_down), sw = ww + 2 * int64_t(params->hs_across);
        static const char *what[8] = {
            "reference window start pixel out of range in down", "reference window start pixel out of range in across",
            "reference window end pixel out of range in down", "reference window end pixel out of range in across",
            "secondary window start pixel out of range in down", "secondary window start pixel out of range in across",
            "secondary window end pixel out of range in down", "secondary window end pixel out of range in across"};
        for (int32_t i = 0; i < n; ++i) {
            const bool fail[8] = {ref_d[i] < 0, ref_a[i] < 0, ref_d[i] + wh >= ref_h, ref_a[i] + ww >= ref_w,
                                  sec_d[i] < 0, sec_a[i] < 0, sec_d[i] + sh >= sec_h, sec_a[i] + sw >= sec_w};
            for (int k = 0; k < 8; ++k)
                if (fail[k]) {
                    if (bad_window) *bad_window = i;
                    if (which) *which = k;
                    throw ampcor::Error(AMPCOR_E_RANGE, std::string(what[k]) + ", window " + std::to_string(i));
                }
        }
    });
}

}  // extern "C"
