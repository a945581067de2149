// CPU restatement of ROI_PAC ampcor (components/mroipac/ampcor/src/ampcor.F),
// the repository's CPU offset tracker, used only as the reported CPU baseline
// (and as a sanity cross-check): per window, time-domain normalized
// correlation (correlate, ampcor.F:1458-1954), SNR gate (:823-843), power-of-two
// FFT x2 oversampling of the complex chips with quadrant spreading (:855-1050,
// fourn2d :2005-2034), correlation of the detected oversampled chips
// (:1072-1079), 8-tap windowed-sinc oversampling of a 16x16 surface patch
// (:1169-1256, fill_sinc/sinc_coef :2270-2438) and the peak search (:1329-1374).
// Arithmetic is single precision like the Fortran (real*4 / complex). The FFT
// is this file's own radix-2 transform (the reference calls FFTW through
// cfft1d_jpl). Windows run in parallel with OpenMP, like DenseAmpcor's
// process-per-row-block split (DenseAmpcor.py:403-436).
//
// Test infrastructure: built by oracle/build_cpu.py into oracle/libampcor_cpu.so
// and called only by the CPU-baseline leg of bench.py and by tests.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

using cpx = std::complex<float>;

// ---------------------------------------------------------------------------
// radix-2 FFT (n a power of two), sign -1 forward, +1 inverse, no scaling
// ---------------------------------------------------------------------------
struct Fft {
    int n = 0, logn = 0;
    std::vector<cpx> w;
    std::vector<int> rev;
    explicit Fft(int n_) : n(n_) {
        while ((1 << logn) < n) ++logn;
        w.resize(n / 2);
        for (int k = 0; k < n / 2; ++k) {
            const double a = -2.0 * M_PI * k / n;
            w[k] = cpx(float(std::cos(a)), float(std::sin(a)));
        }
        rev.resize(n);
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
            rev[i] = r;
        }
    }
    void run(cpx *x, int stride, int sign) const {
        for (int i = 0; i < n; ++i)
            if (i < rev[i]) std::swap(x[i * stride], x[rev[i] * stride]);
        for (int len = 2; len <= n; len <<= 1) {
            const int half = len / 2, step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < half; ++k) {
                    cpx t = w[k * step];
                    if (sign > 0) t = std::conj(t);
                    const cpx a = x[(i + k) * stride], b = x[(i + k + half) * stride] * t;
                    x[(i + k) * stride] = a + b;
                    x[(i + k + half) * stride] = a - b;
                }
        }
    }
};

struct FftCache {
    std::vector<Fft *> by_log;
    FftCache() : by_log(16, nullptr) {}
    ~FftCache() { for (auto *f : by_log) delete f; }
    const Fft &get(int n) {
        int l = 0;
        while ((1 << l) < n) ++l;
        if (!by_log[l]) by_log[l] = new Fft(n);
        return *by_log[l];
    }
};

// 2-D transform of an ny x nx row-major array (fourn2d: rows then columns)
void fft2d(FftCache &fc, cpx *a, int ny, int nx, int sign) {
    const Fft &fx = fc.get(nx), &fy = fc.get(ny);
    for (int r = 0; r < ny; ++r) fx.run(a + size_t(r) * nx, 1, sign);
    for (int c = 0; c < nx; ++c) fy.run(a + c, nx, sign);
}

int nextpow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// ---------------------------------------------------------------------------
// correlate (ampcor.F:1458-1954) with i_avg = 1. Images are [row=down][col=across].
// Returns the normalized surface corn[dn][ac] (lags 0..(sy-ty), 0..(sx-tx)),
// the peak and its lags, covariance and flags.
// ---------------------------------------------------------------------------
struct CorrOut {
    std::vector<float> corn;   // [ny][nx] with ny = sy-ty+1, nx = sx-tx+1
    int nx = 0, ny = 0;
    float peak = 0.f, cov[3] = {0, 0, 0};
    int shx = 0, shy = 0;      // across, down lag of the peak
    int edge_x = 0, edge_y = 0, flag = 1;
};

void correlate(const float *ti, int tx, int ty, const float *sj, int sx, int sy, CorrOut &o) {
    o.nx = sx - tx + 1;
    o.ny = sy - ty + 1;
    o.corn.assign(size_t(o.nx) * o.ny, 0.f);
    o.flag = 1; o.edge_x = o.edge_y = 0;
    o.cov[0] = o.cov[1] = o.cov[2] = 0.f;
    // statistics over non-zero template pixels
    int cnti = 0, cntj = 0;
    float sumi = 0.f, smqi = 0.f;
    for (int y = 0; y < ty; ++y)
        for (int x = 0; x < tx; ++x) {
            const float v = ti[y * tx + x];
            if (v != 0.f) { ++cnti; sumi += v; smqi += v * v; }
        }
    for (int i = 0; i < sx * sy; ++i)
        if (sj[i] != 0.f) ++cntj;
    float meani = 0.f, stdvi = 0.f;
    if (cnti) {
        meani = sumi / cnti;
        stdvi = std::sqrt(smqi / cnti - meani * meani);
    }
    const int nwi = tx * ty;
    if (!(cnti >= 0.9f * nwi && cntj >= 0.9f * sx * sy)) { o.peak = 0.f; return; }
    // window sums of the search image over every lag (running sums)
    std::vector<float> sumj(size_t(o.nx) * o.ny), smqj(size_t(o.nx) * o.ny);
    std::vector<float> rs(size_t(sy) * o.nx), rq(size_t(sy) * o.nx);
    for (int y = 0; y < sy; ++y) {
        float s = 0.f, q = 0.f;
        for (int x = 0; x < tx; ++x) { const float v = sj[y * sx + x]; s += v; q += v * v; }
        rs[y * o.nx] = s; rq[y * o.nx] = q;
        for (int m = 1; m < o.nx; ++m) {
            const float a = sj[y * sx + m - 1], b = sj[y * sx + m + tx - 1];
            s += b - a; q += b * b - a * a;
            rs[y * o.nx + m] = s; rq[y * o.nx + m] = q;
        }
    }
    for (int m = 0; m < o.nx; ++m) {
        float s = 0.f, q = 0.f;
        for (int y = 0; y < ty; ++y) { s += rs[y * o.nx + m]; q += rq[y * o.nx + m]; }
        sumj[m] = s; smqj[m] = q;
        for (int n = 1; n < o.ny; ++n) {
            s += rs[(n + ty - 1) * o.nx + m] - rs[(n - 1) * o.nx + m];
            q += rq[(n + ty - 1) * o.nx + m] - rq[(n - 1) * o.nx + m];
            sumj[n * o.nx + m] = s; smqj[n * o.nx + m] = q;
        }
    }
    // direct correlation and the peak: lag loop across-major as in the Fortran (m outer, n inner)
    float peak = -9.e27f;
    int px = 0, py = 0;
    std::vector<float> crpd(size_t(o.nx) * o.ny, 0.f);
    for (int n = 0; n < o.ny; ++n)
        for (int y = 0; y < ty; ++y) {
            const float *tr = ti + y * tx;
            const float *sr = sj + (n + y) * sx;
            float *out = crpd.data() + size_t(n) * o.nx;
            for (int m = 0; m < o.nx; ++m) {
                float acc = 0.f;
                const float *s = sr + m;
#pragma omp simd reduction(+ : acc)
                for (int x = 0; x < tx; ++x) acc += tr[x] * s[x];
                out[m] += acc;
            }
        }
    for (int m = 0; m < o.nx; ++m)
        for (int n = 0; n < o.ny; ++n) {
            const int k = n * o.nx + m;
            const float corr = crpd[k] - meani * sumj[k];
            const float den = stdvi * std::sqrt(smqj[k] * nwi - sumj[k] * sumj[k]);
            const float c = den > 0.f ? corr / den : 0.f;
            o.corn[k] = c;
            if (peak < c) { peak = c; px = m; py = n; }
        }
    o.peak = peak;
    o.shx = px; o.shy = py;
    if (!(peak > 0.f)) return;
    const int lx = o.nx - 1, ly = o.ny - 1;
    if (py == 0 || py == ly) o.edge_y = 1;
    if (px == 0 || px == lx) o.edge_x = 1;
    auto C = [&](int x, int y) { return o.corn[size_t(y) * o.nx + x]; };
    // curvature at the peak with the reference's one-sided edge forms (ampcor.F:1735-1834)
    float dxx, dyy, dxy, pk = peak;
    const int ix = px, iy = py;
    auto edge_scale = [&](float f) { dxx /= f; dyy /= f; dxy /= f; pk /= f; };
    if (ix == 0 || ix == lx) {
        const int xn = ix == 0 ? ix + 1 : ix - 1;
        dxx = -(C(xn, iy) + C(xn, iy) - 2 * C(ix, iy));
        if (iy == 0 || iy == ly) {
            const int yn = iy == 0 ? iy + 1 : iy - 1;
            dyy = -(C(ix, yn) + C(ix, yn) - 2 * C(ix, iy));
            dxy = 0.f;
            edge_scale(4.f);
        } else {
            dyy = -(C(ix, iy + 1) + C(ix, iy - 1) - 2 * C(ix, iy));
            dxy = ix == 0 ? 2 * (C(ix + 1, iy + 1) - C(ix + 1, iy - 1)) / 4.f
                          : 2 * (C(ix - 1, iy - 1) - C(ix - 1, iy + 1)) / 4.f;
            edge_scale(2.f);
        }
    } else if (iy == 0 || iy == ly) {
        const int yn = iy == 0 ? iy + 1 : iy - 1;
        dxx = -(C(ix + 1, iy) + C(ix - 1, iy) - 2 * C(ix, iy));
        dyy = -(C(ix, yn) + C(ix, yn) - 2 * C(ix, iy));
        dxy = iy == 0 ? 2 * (C(ix + 1, iy + 1) - C(ix - 1, iy + 1)) / 4.f
                      : 2 * (C(ix - 1, iy - 1) - C(ix + 1, iy - 1)) / 4.f;
        edge_scale(2.f);
    } else {
        dxx = -(C(ix + 1, iy) + C(ix - 1, iy) - 2 * C(ix, iy));
        dyy = -(C(ix, iy + 1) + C(ix, iy - 1) - 2 * C(ix, iy));
        dxy = (C(ix + 1, iy + 1) + C(ix - 1, iy - 1) - C(ix + 1, iy - 1) - C(ix - 1, iy + 1)) / 4.f;
    }
    o.peak = pk;
    float n2 = std::max(1.f - pk, 0.f);
    dxx *= nwi; dyy *= nwi; dxy *= nwi;
    float n4 = n2 * n2;
    n2 *= 2.f;
    n4 *= 0.5f * nwi;
    const float u = dxy * dxy - dxx * dyy, u2 = u * u;
    o.flag = 0;
    if (u == 0.f) {
        o.cov[0] = o.cov[1] = 99.f; o.cov[2] = 0.f;
        o.flag = 1;
    } else {
        o.cov[0] = (-n2 * u * dyy + n4 * (dyy * dyy + dxy * dxy)) / u2;
        o.cov[1] = (-n2 * u * dxx + n4 * (dxx * dxx + dxy * dxy)) / u2;
        o.cov[2] = ((n2 * u - n4 * (dxx + dyy)) * dxy) / u2;
    }
}

// derampc (ampcor.F:1958-2001), indices 1-based in the phase as in the Fortran
void derampc(cpx *img, int nx, int ny) {
    cpx phac(0, 0), phdn(0, 0);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx - 1; ++x) phac += img[y * nx + x] * std::conj(img[y * nx + x + 1]);
    for (int y = 0; y < ny - 1; ++y)
        for (int x = 0; x < nx; ++x) phdn += img[y * nx + x] * std::conj(img[(y + 1) * nx + x]);
    const float rac = std::abs(phac) == 0.f ? 0.f : std::atan2(phac.imag(), phac.real());
    const float rdn = std::abs(phdn) == 0.f ? 0.f : std::atan2(phdn.imag(), phdn.real());
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const float ph = rac * (x + 1) + rdn * (y + 1);
            img[y * nx + x] *= cpx(std::cos(ph), std::sin(ph));
        }
}

// x2 spectral spread of an ny x nx spectrum into 2ny x 2nx (ampcor.F:948-970)
void spread2(const cpx *in, int ny, int nx, cpx *out) {
    const int NY = 2 * ny, NX = 2 * nx;
    std::fill(out, out + size_t(NY) * NX, cpx(0, 0));
    for (int k = 0; k < ny / 2; ++k)
        for (int l = 0; l < nx / 2; ++l) {
            out[k * NX + l] = in[k * nx + l];
            out[(NY - ny / 2 + k) * NX + l] = in[(k + ny / 2) * nx + l];
            out[k * NX + NX - nx / 2 + l] = in[k * nx + nx / 2 + l];
            out[(NY - ny / 2 + k) * NX + NX - nx / 2 + l] = in[(k + ny / 2) * nx + nx / 2 + l];
        }
}

// fill_sinc / sinc_coef (ampcor.F:2270-2438): r_fintp(i + j*intplength)
struct Sinc {
    int intplength = 8, decfactor = 4096;
    float fdelay = 4.f;
    std::vector<float> fintp;
    Sinc() {
        const double beta = 0.75, relfiltlen = 6.0, pedestal = 0.0, pi = 4.0 * std::atan(1.0);
        intplength = int(std::lround(relfiltlen / beta));
        const int ncoef = intplength * decfactor;
        const double wgthgt = (1.0 - pedestal) / 2.0, soff = (ncoef - 1.0) / 2.0;
        std::vector<double> filt(ncoef);
        for (int i = 0; i < ncoef; ++i) {
            const double wa = i - soff;
            const double wgt = (1.0 - wgthgt) + wgthgt * std::cos(pi * wa / soff);
            const double s = wa * beta / double(decfactor);
            const double fct = s != 0.0 ? std::sin(pi * s) / (pi * s) : 1.0;
            filt[i] = fct * wgt;
        }
        fdelay = float(intplength / 2.0);
        fintp.assign(size_t(intplength) * decfactor, 0.f);
        for (int i = 0; i < intplength; ++i)
            for (int j = 0; j < decfactor; ++j) fintp[i + j * intplength] = float(filt[j + i * decfactor]);
    }
};

struct Params {
    int W, hs, srchp, cw, covs, sinc_window;
    float snrth, covth;
};

struct Result { float off_down, off_across, snr, cov[3]; int ok; };

Result one_window(FftCache &fc, const Sinc &sk, const Params &p, const float *ref, int64_t ref_pitch,
                  const float *sec, int64_t sec_pitch, int64_t rd, int64_t ra, int64_t sd, int64_t sa) {
    Result res{0.f, 0.f, 0.f, {99.f, 99.f, 0.f}, 0};
    const int W = p.W, S = W + 2 * p.hs;
    auto at = [](const float *img, int64_t pitch, int64_t r, int64_t c) {
        return cpx(img[2 * (r * pitch + c)], img[2 * (r * pitch + c) + 1]);
    };
    std::vector<float> ti(W * W), sj(S * S);
    for (int y = 0; y < W; ++y)
        for (int x = 0; x < W; ++x) ti[y * W + x] = std::abs(at(ref, ref_pitch, rd + y, ra + x));
    for (int y = 0; y < S; ++y)
        for (int x = 0; x < S; ++x) sj[y * S + x] = std::abs(at(sec, sec_pitch, sd + y, sa + x));
    CorrOut c1;
    correlate(ti.data(), W, W, sj.data(), S, S, c1);
    if (c1.flag || c1.edge_x || c1.edge_y) return res;
    // SNR over the peak neighbourhood (-9..+11 lags, ampcor.F:823-838)
    float outside = 0.f;
    int cnt = 0;
    for (int l = std::max(c1.shy - 9, 1); l <= std::min(c1.shy + 11, S - W); ++l)
        for (int k = std::max(c1.shx - 9, 1); k <= std::min(c1.shx + 11, S - W); ++k) {
            const float v = c1.corn[size_t(l - 1) * c1.nx + (k - 1)];
            outside += v * v;
            ++cnt;
        }
    outside = (outside - c1.peak * c1.peak) / (cnt - 1);
    const float snr = c1.peak * c1.peak / std::max(outside, 1.e-10f);
    if (!(snr > p.snrth && c1.cov[0] < p.covth && c1.cov[1] < p.covth)) return res;
    const float shx = float(c1.shx) - p.hs, shy = float(c1.shy) - p.hs;
    // oversample the reference chip and the search chip around the peak
    const int nri = nextpow2(W), wjp = W + 2 * p.srchp, nrj = nextpow2(wjp);
    std::vector<cpx> cref(size_t(nri) * nri, cpx(0, 0)), csch(size_t(nrj) * nrj, cpx(0, 0));
    for (int y = 0; y < W; ++y)
        for (int x = 0; x < W; ++x) cref[y * nri + x] = at(ref, ref_pitch, rd + y, ra + x);
    const int64_t oy = sd + c1.shy - p.srchp, ox = sa + c1.shx - p.srchp;
    for (int y = 0; y < wjp; ++y)
        for (int x = 0; x < wjp; ++x) csch[y * nrj + x] = at(sec, sec_pitch, oy + y, ox + x);
    derampc(cref.data(), nri, nri);
    derampc(csch.data(), nrj, nrj);
    fft2d(fc, csch.data(), nrj, nrj, -1);
    fft2d(fc, cref.data(), nri, nri, -1);
    std::vector<cpx> osref(size_t(4) * nri * nri), ossch(size_t(4) * nrj * nrj);
    spread2(cref.data(), nri, nri, osref.data());
    spread2(csch.data(), nrj, nrj, ossch.data());
    fft2d(fc, osref.data(), 2 * nri, 2 * nri, +1);
    fft2d(fc, ossch.data(), 2 * nrj, 2 * nrj, +1);
    const int Wo = 2 * W, Jo = 2 * wjp;
    std::vector<float> tio(Wo * Wo), sjo(Jo * Jo);
    const float si = 1.f / float(nri * nri), sjs = 1.f / float(nrj * nrj);
    for (int y = 0; y < Wo; ++y)
        for (int x = 0; x < Wo; ++x) tio[y * Wo + x] = std::abs(osref[size_t(y) * 2 * nri + x] * si);
    for (int y = 0; y < Jo; ++y)
        for (int x = 0; x < Jo; ++x) sjo[y * Jo + x] = std::abs(ossch[size_t(y) * 2 * nrj + x]) * sjs;
    CorrOut c2;
    correlate(tio.data(), Wo, Wo, sjo.data(), Jo, Jo, c2);
    const int wso = Jo - (Wo - 1);
    const float shxos = float(c2.shx) / 2.f - float((wso - 1) / 2) / 2.f + shx;
    const float shyos = float(c2.shy) / 2.f - float((wso - 1) / 2) / 2.f + shy;
    // 16x16 patch of |surface|/peak around the peak, zero outside (ampcor.F:1094-1120)
    const int cw = p.cw, covs = p.covs;
    std::vector<float> corrt(cw * cw, 0.f);
    const int ppx = c2.shx + 1, ppy = c2.shy + 1;
    for (int yy = -cw / 2; yy < cw / 2; ++yy)
        for (int xx = -cw / 2; xx < cw / 2; ++xx) {
            const int idx = (yy + cw / 2) * cw + xx + cw / 2;
            // the Fortran surface array is zero beyond the computed lags
            if (xx + ppx >= 1 && xx + ppx <= (2 * p.srchp + 1) * 2 && yy + ppy >= 1 &&
                yy + ppy <= (2 * p.srchp + 1) * 2 && xx + ppx <= c2.nx && yy + ppy <= c2.ny)
                corrt[idx] = std::fabs(c2.corn[size_t(yy + ppy - 1) * c2.nx + (xx + ppx - 1)] / c2.peak);
        }
    // two 1-D sinc passes over +-sinc_window*covs around the centre (ampcor.F:1169-1256)
    const int N = cw * covs, R = p.sinc_window * covs;
    std::vector<float> d1(size_t(cw) * N, 0.f), d2(size_t(N) * N, 0.f);
    for (int yy = -cw / 2; yy < cw / 2; ++yy)
        for (int xx = -R; xx <= R; ++xx) {
            const float rj = float(xx + N / 2 + covs) / covs + sk.fdelay;
            const int ij = int(rj);
            const int fr = int((rj - ij) * sk.decfactor);
            float acc = 0.f, wsum = 0.f;
            for (int k = 0; k < sk.intplength; ++k) {
                int jj = ij - k;
                if (jj < 1) jj += cw; else if (jj > cw) jj -= cw;
                const float wt = sk.fintp[k + fr * sk.intplength];
                acc += corrt[(yy + cw / 2) * cw + jj - 1] * wt;
                wsum += wt;
            }
            d1[size_t(yy + cw / 2) * N + xx + N / 2] = acc / wsum;
        }
    for (int yy = -R; yy <= R; ++yy)
        for (int xx = -R; xx <= R; ++xx) {
            const float ri = float(yy + N / 2 + covs) / covs + sk.fdelay;
            const int ii = int(ri);
            const int fr = int((ri - ii) * sk.decfactor);
            float acc = 0.f, wsum = 0.f;
            for (int k = 0; k < sk.intplength; ++k) {
                int jj = ii - k;
                if (jj < 1) jj += cw; else if (jj > cw) jj -= cw;
                const float wt = sk.fintp[k + fr * sk.intplength];
                acc += d1[size_t(jj - 1) * N + xx + N / 2] * wt;
                wsum += wt;
            }
            d2[size_t(yy + N / 2) * N + xx + N / 2] = acc / wsum;
        }
    // peak within +-covs of the centre; later equal values win (ampcor.F:1329-1347)
    float rmax = 0.f;
    int cpx_ = 0, cpy_ = 0;
    for (int y = 1; y <= N; ++y)
        for (int x = 1; x <= N; ++x) {
            const float v = std::fabs(d2[size_t(y - 1) * N + x - 1]);
            if (std::abs(x - N / 2) <= covs && std::abs(y - N / 2) <= covs && v >= rmax) {
                rmax = v;
                cpx_ = x - cw / 2 * covs;
                cpy_ = y - cw / 2 * covs;
            }
        }
    const float offx = float(cpx_ - 1) / float(covs), offy = float(cpy_ - 1) / float(covs);
    res.off_across = offx / 2.f + shxos;
    res.off_down = offy / 2.f + shyos;
    res.snr = std::min(snr, 9999.99999f);
    res.cov[0] = c1.cov[0]; res.cov[1] = c1.cov[1]; res.cov[2] = c1.cov[2];
    res.ok = 1;
    return res;
}

}  // namespace

extern "C" {

// Run ampcor on n windows; starts are 0-based (reference chip and search chip
// origins, the search chip being (W+2hs)^2). Images are complex64 interleaved.
// Outputs: off[n][2] (down, across; NaN where the reference rejects the
// window), snr[n], cov[n][3]. Returns the number of windows accepted.
int ampcor_cpu_run(const float *ref, int64_t ref_pitch, const float *sec, int64_t sec_pitch,
                   const int32_t *ref_d, const int32_t *ref_a, const int32_t *sec_d, const int32_t *sec_a,
                   int32_t n, int32_t W, int32_t hs, int32_t nthreads, float *off, float *snr, float *cov) {
    Params p{W, hs, std::min(hs, 4), 16, 16, 2, 1.0e-6f, 999.999998f};
    const Sinc sk;
    int accepted = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : accepted)
    {
        FftCache fc;
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < n; ++i) {
            const Result r = one_window(fc, sk, p, ref, ref_pitch, sec, sec_pitch, ref_d[i], ref_a[i], sec_d[i],
                                        sec_a[i]);
            off[2 * i] = r.ok ? r.off_down : NAN;
            off[2 * i + 1] = r.ok ? r.off_across : NAN;
            snr[i] = r.snr;
            cov[3 * i] = r.cov[0]; cov[3 * i + 1] = r.cov[1]; cov[3 * i + 2] = r.cov[2];
            accepted += r.ok;
        }
    }
    return accepted;
}

int ampcor_cpu_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
