"""Pin the FP64 oracle (oracle/ampcor_oracle.py) with checks that do not restate it.

The reference ships no tests, fixtures or golden vectors for PyCuAmpcor, so each
stage is pinned by brute force, closed forms, emulation of the reference's
thread structure, or invariants a plausible mistake (sign, index, transpose,
dropped term) would break.
"""
import numpy as np
import pytest

from oracle import ampcor_oracle as O

rng = np.random.default_rng(1234)


# ---------------------------------------------------------------------------
# correlation and normalization
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("th,tw,sh,sw", [(4, 5, 9, 11), (8, 8, 14, 14), (3, 7, 10, 8)])
def test_correlate_matches_brute_force(th, tw, sh, sw):
    t = rng.standard_normal((3, th, tw))
    s = rng.standard_normal((3, sh, sw))
    a = O.correlate(t, s)
    b = O.correlate_direct(t, s)
    assert a.shape == (3, sh - th + 1, sw - tw + 1)
    np.testing.assert_allclose(a, b, atol=1e-10)


def test_correlate_integer_shift_peak():
    s = rng.random((1, 40, 40))
    u, v = 7, 3
    t = s[:, u:u + 16, v:v + 16].copy()
    c = O.normalize(O.correlate(O.subtract_mean(t), s), O.subtract_mean(t), s)
    loc, val = O.maxloc2d(c)
    assert tuple(loc[0]) == (u, v)
    assert abs(val[0] - 1.0) < 1e-6


def test_window_sums_brute_force():
    s = rng.random((2, 13, 17))
    s1, s2 = O.window_sums(s, 5, 6)
    for u in range(13 - 5 + 1):
        for v in range(17 - 6 + 1):
            blk = s[:, u:u + 5, v:v + 6]
            np.testing.assert_allclose(s1[:, u, v], blk.sum(axis=(1, 2)), rtol=1e-12)
            np.testing.assert_allclose(s2[:, u, v], (blk ** 2).sum(axis=(1, 2)), rtol=1e-12)


def test_normalize_invariances():
    t = O.subtract_mean(rng.random((2, 8, 8)) + 1)
    s = rng.random((2, 20, 20)) + 1
    c = O.normalize(O.correlate(t, s), t, s)
    # scale invariance up to the FLT_EPSILON regularizer (negligible at these magnitudes)
    c2 = O.normalize(O.correlate(3 * t, 5 * s), 3 * t, 5 * s)
    np.testing.assert_allclose(c, c2, rtol=1e-6)
    # offset invariance of the search (mean-subtracted template)
    c3 = O.normalize(O.correlate(t, s + 7), t, s + 7)
    np.testing.assert_allclose(c, c3, rtol=1e-6, atol=1e-9)
    assert np.all(np.abs(c) <= 1 + 1e-9)


def test_normalize_flt_epsilon_regularizer():
    # a constant search window has zero variance: the result is corr / sqrt(FLT_EPSILON)
    t = O.subtract_mean(rng.random((1, 4, 4)))
    s = np.full((1, 6, 6), 2.0)
    c = O.correlate(t, s)
    n = O.normalize(c, t, s)
    np.testing.assert_allclose(n, c / np.sqrt(O.FLT_EPSILON), atol=1e-12)


# ---------------------------------------------------------------------------
# argmax tie rule: emulate the reference's 512-thread scan + strided tree
# ---------------------------------------------------------------------------
def _maxloc_threads(surface):
    flat = surface.ravel()
    B = 512
    val = [-O.FLT_MAX] * B
    loc = [None] * B
    for tid in range(B):
        for gid in range(tid, flat.size, B):
            if val[tid] < flat[gid]:
                val[tid], loc[tid] = flat[gid], gid
    s = B // 2
    while s >= 1:
        for tid in range(s):
            if val[tid] < val[tid + s]:
                val[tid], loc[tid] = val[tid + s], loc[tid + s]
        s //= 2
    return loc[0], val[0]


@pytest.mark.parametrize("shape", [(41, 41), (17, 17), (33, 33), (25, 40)])
def test_maxloc_tie_rule_matches_thread_emulation(shape):
    for trial in range(6):
        surf = rng.integers(0, 4, size=shape).astype(float)  # many exact ties
        loc, val = O.maxloc2d(surf[None])
        gid, v = _maxloc_threads(surf)
        assert loc[0, 0] * shape[1] + loc[0, 1] == gid
        assert val[0] == v


def test_maxloc_all_equal_and_nan():
    loc, _ = O.maxloc2d(np.zeros((1, 41, 41)))
    assert tuple(loc[0]) == (0, 0)
    s = rng.random((1, 9, 9))
    s[0, 4, 4] = np.nan
    s[0, 2, 3] = 5.0
    loc, val = O.maxloc2d(s)
    assert tuple(loc[0]) == (2, 3) and val[0] == 5.0


# ---------------------------------------------------------------------------
# covariance, SNR, zoom start, sub-pixel
# ---------------------------------------------------------------------------
def test_covariance_quadratic_closed_form():
    # c = p - a x^2 - b y^2 around the peak: dxx = 2aT, dyy = 2bT, dxy = 0, and the
    # reference formula reduces to cov = (1-p)/(a T) + (1-p)^2/(8 a^2 T)
    p, a, b, T = 0.8, 0.05, 0.03, 64 * 64
    x = np.arange(41)[:, None] - 20
    y = np.arange(41)[None, :] - 20
    surf = (p - a * x ** 2 - b * y ** 2)[None]
    cov = O.estimate_variance(surf, np.array([[20, 20]]), np.array([p]), T)[0]
    np.testing.assert_allclose(cov[0], (1 - p) / (a * T) + (1 - p) ** 2 / (8 * a * a * T), rtol=1e-12)
    np.testing.assert_allclose(cov[1], (1 - p) / (b * T) + (1 - p) ** 2 / (8 * b * b * T), rtol=1e-12)
    assert abs(cov[2]) < 1e-15
    # swapping the axes swaps the variances and keeps the cross term
    cov_t = O.estimate_variance(surf.transpose(0, 2, 1), np.array([[20, 20]]), np.array([p]), T)[0]
    np.testing.assert_allclose(cov_t[:2], cov[1::-1], rtol=1e-12)


def test_covariance_sentinels():
    surf = rng.random((2, 41, 41))
    cov = O.estimate_variance(surf, np.array([[0, 5], [40, 40]]), np.array([1.0, 1.0]), 4096)
    np.testing.assert_array_equal(cov, [[99, 99, 0], [99, 99, 0]])
    flat = np.full((1, 41, 41), 0.5)   # zero curvature: |u| < 1e-2
    cov = O.estimate_variance(flat, np.array([[20, 20]]), np.array([0.5]), 4096)
    np.testing.assert_array_equal(cov[0], [99, 99, 0])


def test_snr_closed_form():
    b, p = 0.1, 0.9
    surf = np.full((1, 41, 41), b)
    surf[0, 20, 20] = p
    snr = O.estimate_snr(surf, np.array([[20, 20]]), np.array([p]), 21, 21)
    np.testing.assert_allclose(snr, p * p / (b * b), rtol=1e-12)
    # at a corner only 11x11 = 121 values are valid
    surf2 = np.full((1, 41, 41), b)
    surf2[0, 0, 0] = p
    snr2 = O.estimate_snr(surf2, np.array([[0, 0]]), np.array([p]), 21, 21)
    np.testing.assert_allclose(snr2, p * p / b ** 2, rtol=1e-12)


def test_adjust_offset_invariants():
    hs, hz = 20, 4
    m = np.arange(0, 2 * hs + 1)
    start, shift = O.adjust_offset(m, hs, hz)
    assert np.all((start >= 0) & (start <= 2 * (hs - hz)))
    inner = (m >= hz) & (m <= 2 * hs - hz)
    assert np.all(start[inner] == m[inner] - hz) and np.all(shift[inner] == 0)
    # edge windows: shift is |overflow| on both sides (cuOffset.cu:211-222)
    assert np.all(shift[~inner] == np.abs((m - hz) - start)[~inner])
    assert np.all(shift >= 0)


def test_subpixel_offset_formula():
    off = O.subpixel_offset(np.array([[16, 3]]), np.array([[128, 40]]), 16, 2, 20, 20)
    np.testing.assert_array_equal(off, [[16 + 4 - 20, 3 + 40 / 32 - 20]])


def test_parabolic_vertex_closed_form():
    """A parabola sampled at -1, 0, 1 gives back its vertex exactly; values that
    do not bend down give 0; the result is clamped to [-1/2, 1/2]."""
    x0 = rng.uniform(-0.5, 0.5, 50)
    a = rng.uniform(0.1, 3.0, 50)
    c = rng.uniform(-1, 1, 50)
    f = lambda x: c - a * (x - x0) ** 2  # noqa: E731
    np.testing.assert_allclose(O.parabolic_vertex(f(-1.0), f(0.0), f(1.0)), x0, atol=1e-12)
    assert O.parabolic_vertex(1.0, 1.0, 1.0) == 0.0 and O.parabolic_vertex(2.0, 1.0, 2.0) == 0.0
    assert O.parabolic_vertex(0.0, 0.1, 5.0) == 0.0                 # bends up
    assert O.parabolic_vertex(0.0, 1.0, 1.5) == 0.5                 # vertex at 1.5, clamped
    assert O.parabolic_vertex(1.5, 1.0, 0.0) == -0.5


def test_refine_offset_quadratic_surface():
    """A separable quadratic peak anywhere on the 256x256 grid (an edge peak
    wrapping around) is located exactly."""
    X = np.arange(256)[:, None]
    Y = np.arange(256)[None, :]
    for x0, y0 in [(100.3, 57.8), (0.2, 255.4), (255.45, 3.0)]:
        dx = (X - x0 + 128) % 256 - 128          # periodic distance
        dy = (Y - y0 + 128) % 256 - 128
        s = (1.0 - 0.003 * dx ** 2 - 0.002 * dy ** 2)[None]
        zidx, _ = O.maxloc2d(s)
        r = O.refine_offset(s, zidx, np.array([[3, 5]]), 16, 2, 4, 4)
        want_x = (x0 if x0 < 255.5 else x0 - 256) / 32 + 3 - 4
        want_y = (y0 if y0 < 255.5 else y0 - 256) / 32 + 5 - 4
        # an edge peak's vertex can lie half a cell beyond the grid: compare modulo the side
        gx = (r[0, 0] - 3 + 4) * 32
        gy = (r[0, 1] - 5 + 4) * 32
        assert min(abs(gx - x0), abs(gx - x0 + 256), abs(gx - x0 - 256)) < 1e-9, (r, want_x)
        assert min(abs(gy - y0), abs(gy - y0 + 256), abs(gy - y0 - 256)) < 1e-9, (r, want_y)


def test_refine_offset_band_limited_peak():
    """A 16x16 sampled 2-D cosine is band-limited, so its FFT-oversampled
    surface is the continuous cosine: the refined peak lands on the cosine's
    maximum (input-sample position times 16) within 1e-3 of a fine cell,
    where the bare argmax is off by up to half a cell."""
    i = np.arange(16)
    for x0, y0 in [(7.23, 8.61), (3.05, 12.97), (10.5, 0.49)]:
        a = (np.cos(2 * np.pi * (i[:, None] - x0) / 16) + np.cos(2 * np.pi * (i[None, :] - y0) / 16))[None]
        surf = O.oversample_r2r(a, 256, 256)
        zidx, _ = O.maxloc2d(surf)
        r = O.refine_offset(surf, zidx, np.zeros((1, 2), int), 16, 1, 0, 0)
        assert abs(r[0, 0] * 16 - x0 * 16) < 1e-3 and abs(r[0, 1] * 16 - y0 * 16) < 1e-3, (r * 16, x0 * 16, y0 * 16)


# ---------------------------------------------------------------------------
# oversampling and deramp
# ---------------------------------------------------------------------------
def _dft_matrix(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


@pytest.mark.parametrize("h,w,H,W", [(6, 8, 12, 16), (5, 7, 10, 14), (4, 4, 64, 64)])
def test_oversample_c2c_brute_force(h, w, H, W):
    x = rng.standard_normal((2, h, w)) + 1j * rng.standard_normal((2, h, w))
    # explicit matrices: +i transform, quadrant placement scaled 1/(h w), -i transform
    X = np.einsum("ab,nbc,cd->nad", _dft_matrix(h, +1), x, _dft_matrix(w, +1))
    P = np.zeros((2, H, W), complex)
    for a in range(h // 2):
        for b in range(w // 2):
            P[:, a, b] = X[:, a, b]
            P[:, H - 1 - a, b] = X[:, h - 1 - a, b]
            P[:, a, W - 1 - b] = X[:, a, w - 1 - b]
            P[:, H - 1 - a, W - 1 - b] = X[:, h - 1 - a, w - 1 - b]
    P /= h * w
    ref = np.einsum("ab,nbc,cd->nad", _dft_matrix(H, -1), P, _dft_matrix(W, -1))
    np.testing.assert_allclose(O.oversample_c2c(x, H, W), ref, atol=1e-10)


def test_oversample_even_samples_equal_input():
    x = rng.standard_normal((3, 16, 12)) + 1j * rng.standard_normal((3, 16, 12))
    y = O.oversample_c2c(x, 32, 24)
    np.testing.assert_allclose(y[:, ::2, ::2], x, atol=1e-12)


def test_oversample_tones_and_nyquist_placement():
    n = 16
    idx = np.arange(n)
    p = np.arange(2 * n)
    for f in range(-n // 2, n // 2):
        x = np.exp(-2j * np.pi * f * idx / n)
        img = np.outer(x, np.ones(4))[None]            # tone along rows, constant across
        y = O.oversample_c2c(img, 2 * n, 8)[0, :, 0]
        np.testing.assert_allclose(y, np.exp(-2j * np.pi * f * p / (2 * n)), atol=1e-10)
    # the Nyquist tone (-1)^n lands on the negative side: upsampled to i^p, not (-i)^p
    img = np.outer((-1.0) ** idx, np.ones(4))[None].astype(complex)
    y = O.oversample_c2c(img, 2 * n, 8)[0, :, 0]
    np.testing.assert_allclose(y, 1j ** p, atol=1e-10)


def test_oversample_r2r_is_real_part():
    s = rng.random((2, 16, 16))
    np.testing.assert_allclose(O.oversample_r2r(s, 256, 256), O.oversample_c2c(s + 0j, 256, 256).real)
    np.testing.assert_allclose(O.oversample_r2r(s, 256, 256)[:, ::16, ::16], s, atol=1e-12)


def test_deramp_operator_quirk_closed_form():
    # symmetric ramp z = exp(i a (r + c)): both phase estimates are -a, the rotation is
    # by -a(r+c); float2's operator*= gives x' = 1, y' = sin(t)(cos(t) - 1), t = a(r+c)
    h = w = 8
    a = 0.3
    r = np.arange(h)[:, None]
    c = np.arange(w)[None, :]
    z = np.exp(1j * a * (r + c))[None]
    out = O.deramp_method1(z)[0]
    t = a * (r + c)
    np.testing.assert_allclose(out.real, 1.0, atol=1e-12)
    np.testing.assert_allclose(out.imag, np.sin(t) * (np.cos(t) - 1), atol=1e-12)


def test_deramp_axes_swapped():
    # ramp only along rows (down): the reference applies its down phase along columns
    h = w = 8
    a = 0.2
    r = np.arange(h)[:, None] + 0 * np.arange(w)[None, :]
    z = np.exp(1j * a * r)[None]
    out = O.deramp_method1(z)[0]
    ph = -a * np.arange(w)[None, :] + 0 * r       # col * phaseDown, phaseDown = -a
    x, y = np.cos(a * r), np.sin(a * r)
    c, s = np.cos(ph), np.sin(ph)
    nx = x * c - y * s
    np.testing.assert_allclose(out.real, nx, atol=1e-12)
    np.testing.assert_allclose(out.imag, y * c + nx * s, atol=1e-12)


def test_sinc_table():
    t, L, dec = O.sinc_filter_table()
    assert L == 8 and dec == 4096 and t.size == 8 * 4096 + 1
    mid = (t.size - 2) / 2          # centre (ncoef-1)/2; the table has one extra tap at the end
    # symmetric about its centre, window zero at the ends, sinc zeros at multiples of dec/beta
    np.testing.assert_allclose(t[:-1], t[-2::-1], atol=1e-12)
    assert t[0] == pytest.approx(0.0, abs=1e-12)
    i = int(round(mid + dec / 0.75))
    x = (i - mid) * 0.75 / dec * O.SINC_PI
    assert abs(np.sin(x) / x) < 1e-3 and abs(t[i]) < 1e-3


def test_sinc_oversample_window_and_identity():
    surf = rng.random((1, 16, 16))
    out = O.oversample_sinc(surf, 256, 16, np.array([[0, 0]]), 32)
    nz = np.argwhere(out[0] != 0)
    assert nz[:, 0].min() == 96 and nz[:, 0].max() == 160
    # on-grid outputs reproduce the input only approximately (8-tap windowed sinc)
    assert np.abs(out[0, 128, 128] - surf[0, 8, 8]) < 0.2


def _tone16(f, g, c1, c2):
    i = np.arange(16)
    return np.cos(2 * np.pi * f * (i[:, None] - c1)) * np.cos(2 * np.pi * g * (i[None, :] - c2))


@pytest.mark.parametrize("covs,factor,shift,f,g", [
    (16, 32, (0, 0), 2 / 16, 3 / 16),
    (16, 32, (3, 5), 2 / 16, 3 / 16),     # shifted window that wraps past the 256 edge
    (16, 16, (7, 0), 1 / 16, 2 / 16),
    (8, 16, (2, 1), 3 / 16, 1 / 16),
    (8, 8, (0, 6), 2 / 16, 2 / 16),
])
def test_sinc_oversample_band_limited_tone(covs, factor, shift, f, g):
    """Interpolation meaning: output sample o is the input signal at position
    o/covs. A separable tone periodic over the 16 samples (so the circular tap
    wrap is consistent) must come back at every fractional phase of the window,
    within the 8-tap windowed-sinc passband error (0.011 at 1/16 cycle/sample).
    A reversed tap direction or a mirrored fractional phase misses by ~0.4-1.1."""
    c1, c2 = 0.3, -1.7
    out_n = 16 * covs
    s = _tone16(f, g, c1, c2)[None]
    out = O.oversample_sinc(s, out_n, covs, np.array([shift]), factor)
    rng_ = 2 * covs
    start = out_n // 2 - rng_
    ox = (np.arange(2 * rng_ + 1) + start + shift[0] * factor) % out_n
    oy = (np.arange(2 * rng_ + 1) + start + shift[1] * factor) % out_n
    truth = (np.cos(2 * np.pi * f * (ox[:, None] / covs - c1))
             * np.cos(2 * np.pi * g * (oy[None, :] / covs - c2)))
    assert np.abs(out[0][np.ix_(ox, oy)] - truth).max() < 0.03
    # everything outside the (4 covs + 1)^2 patch stays zero
    mask = np.ones((out_n, out_n), bool)
    mask[np.ix_(ox, oy)] = False
    assert not out[0][mask].any()


def _sinc_point_fp32(surf, outx, outy, covs, table):
    """One output of the reference's interpolation kernel, scalar and in FP32:
    r_xout = outx/covs, i_xfrac = int(frac*4096), taps inx = i_xout - i + 4 with
    the wrap (x wrapped by the input width), weights normalized by their sum
    (cuSincOverSampler.cu:108-157)."""
    f32 = np.float32
    n = surf.shape[0]
    rx = f32(outx) / f32(covs)
    ix = int(rx)
    fx = int((rx - f32(ix)) * f32(4096))
    ry = f32(outy) / f32(covs)
    iy = int(ry)
    fy = int((ry - f32(iy)) * f32(4096))
    acc, wsum = f32(0), f32(0)
    for i in range(8):
        inx = ix - i + 4
        inx = inx + n if inx < 0 else inx
        inx = inx - n if inx >= n else inx
        cx = f32(table[i * 4096 + fx])
        for j in range(8):
            iny = iy - j + 4
            iny = iny + n if iny < 0 else iny
            iny = iny - n if iny >= n else iny
            c = cx * f32(table[j * 4096 + fy])
            wsum += c
            acc += f32(surf[inx, iny]) * c
    return float(acc / wsum)


def test_sinc_oversample_matches_index_formulas():
    surf = rng.random((1, 16, 16))
    table = O.sinc_filter_table()[0]
    shift = np.array([[2, 4]])
    out = O.oversample_sinc(surf, 256, 16, shift, 32)
    for ox, oy in [(96 + 64, 96 + 128), (100 + 64, 131 + 128 - 256), (160 + 64, 97 + 128 - 256),
                   (123 + 64, 111 + 128)]:
        ox %= 256
        oy %= 256
        ref = _sinc_point_fp32(surf[0], ox, oy, 16, table)
        assert out[0, ox, oy] == pytest.approx(ref, rel=2e-6, abs=2e-6)


# ---------------------------------------------------------------------------
# start pixels, range check, whole pipeline
# ---------------------------------------------------------------------------
def test_start_pixels_and_range_check():
    rd, ra, sd, sa, gd, ga = O.start_pixels(2, 3, 30, 20, 64, 32, 5, -3, 20, 16)
    np.testing.assert_array_equal(rd, [30, 30, 30, 94, 94, 94])
    np.testing.assert_array_equal(ra, [20, 52, 84, 20, 52, 84])
    np.testing.assert_array_equal(sd, rd + 5 - 20)
    np.testing.assert_array_equal(sa, ra - 3 - 16)
    cfg = O.AmpcorConfig(hs_down=20, hs_across=16)
    assert O.check_range(rd, ra, sd, sa, cfg, (200, 200), (200, 200)) is None
    # strict end test: start + size == dim - 0 passes only if < dim
    bad = O.check_range(rd, ra, sd, sa, cfg, (200, 148), (200, 200))
    assert bad == (2, "Reference Window end pixel out of range in Across")
    bad = O.check_range(rd - 31, ra, sd, sa, cfg, (200, 200), (200, 200))
    assert bad[0] == 0 and "Down" in bad[1]


def test_derived_sizes_reference_defaults():
    d = O.derived_sizes(O.AmpcorConfig())
    assert (d["search_h"], d["zoom_in_h"], d["os_win_h"], d["os_search_h"]) == (104, 72, 128, 144)
    assert (d["raw_corr_h"], d["os_corr_h"], d["surf_in"], d["surf_out"]) == (41, 17, 16, 256)
    assert d["corr_raw_zoom_h"] == 21 and d["hz"] == 4
    d = O.derived_sizes(O.AmpcorConfig(hs_down=3, hs_across=3))
    assert d["zoom"] == 12 and d["hz"] == 3


def test_pipeline_recovers_known_shift():
    from isce2_amd.synth import speckle_pair_shift
    ref, sec = speckle_pair_shift(256, 256, shift=(3.25, -1.75), snr_db=None, seed=3)
    cfg = O.AmpcorConfig(hs_down=16, hs_across=16)
    r = O.run_ampcor(ref, sec, cfg, 2, 2, 40, 40, 64, 64)
    err = r["offset"] - np.array([3.25, -1.75])
    assert np.all(np.abs(err) <= 1 / 32 + 1e-12), err
    assert np.median(np.abs(err)) <= 1 / 64


@pytest.mark.parametrize("deramp", [0, 1, 2])
def test_pipeline_modes_and_time_domain(deramp):
    from isce2_amd.synth import speckle_pair_shift
    ref, sec = speckle_pair_shift(200, 200, shift=(1.5, 2.0), snr_db=20, seed=5)
    cfg = O.AmpcorConfig(deramp=deramp, hs_down=8, hs_across=8, window_h=32, window_w=32)
    r0 = O.run_ampcor(ref, sec, cfg, 2, 2, 30, 30, 48, 48)
    assert np.all(np.abs(r0["offset"] - [1.5, 2.0]) <= 2 / 32)
    cfg.algorithm = 1   # the time-domain correlator computes the same sums
    r1 = O.run_ampcor(ref, sec, cfg, 2, 2, 30, 30, 48, 48)
    np.testing.assert_array_equal(r0["offset"], r1["offset"])
