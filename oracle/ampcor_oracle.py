"""FP64 restatement of PyCuAmpcor's per-window offset-tracking pipeline.

TEST INFRASTRUCTURE ONLY: this module is the oracle the HIP engine is checked
against. It is imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``; the product package never imports it.

Every stage follows one reference function (paths relative to the isce2 tree,
``contrib/PyCuAmpcor/src/``), restated in float64 NumPy:

=========================  ====================================================
stage                      reference
=========================  ====================================================
derived sizes              cuAmpcorParameter.cpp:70-109
start pixels               cuAmpcorParameter.cpp:155-205
range check                cuAmpcorParameter.cpp:275-331
chip gather (+|z| mode)    cuArraysCopy.cu:70-165, cuAmpcorChunk.cpp:260-372
amplitude                  cuArraysCopy.cu:510-529 (float2.h:114-117)
mean subtraction           cuCorrNormalization.cu:98-139
correlation (freq/time)    cuCorrFrequency.cu:63-111, cuCorrTimeDomain.cu:12-92
normalization              cuCorrNormalization.cu:229-314 (+SAT :203-242)
argmax                     cuOffset.cu:13-84
covariance                 cuEstimateStats.cu:49-110
SNR                        cuArraysCopy.cu:249-284, cuCorrNormalization.cu:143-192,
                           cuEstimateStats.cu:20-29
zoom start / shift         cuOffset.cu:190-241
deramp (method 1)          cuDeramp.cu:67-148 with float2.h:71-75
chip oversampling (C2C)    cuOverSampler.cpp:72-78, cuArraysPadding.cu:59-103
surface oversampling FFT   cuOverSampler.cpp:141-148, cuArraysCopy.cu:326-339
surface oversampling sinc  cuSincOverSampler.cu:21-195, cuSincOverSampler.h:27-57
sub-pixel offset           cuOffset.cu:137-187
chunk run / outputs        cuAmpcorChunk.cpp:12-217, cuAmpcorController.cpp:37-158
=========================  ====================================================

Reference quirks reproduced on purpose (each is what PyCuAmpcor computes):

* Deramp multiplies with ``float2 operator*=`` (float2.h:71-75), which writes
  the real part first and then uses the *new* real part for the imaginary
  part; the phase is ``col*phaseDown + row*phaseAcross`` (axes swapped).
* The chip oversampler transforms with the +i sign first, pads the four
  ``h//2 x w//2`` spectrum quadrants into the corners scaled by ``1/(h*w)``
  (odd sizes drop the middle row/column), then transforms with -i; chips are
  used as ``|z|`` afterwards while the surface oversampler keeps the real part.
* Normalization uses ``rsqrt((sum s^2 - (sum s)^2/|t|) * sum t'^2 + FLT_EPSILON)``
  with the mean-subtracted template and the raw search amplitudes.
* Argmax ties resolve the way the 512-thread strided scan + tree reduction of
  cuOffset.cu does: the smallest bit-reversed ``idx % 512``, then smallest idx.
* The sinc-oversampler centre shift is always >= 0 (cuOffset.cu:211-222 sets
  ``shift = -start`` on the left edge), and the filter table uses the float
  constant ``PI = 3.14159265359``.

One deliberate divergence: for unequal down/across search ranges the
reference sizes the search chip with the swapped half ranges
(cuAmpcorParameter.cpp:90-91), reading outside the chip when the down range
is larger. Here each axis uses its own half range; for equal ranges (every
configuration the reference handles in bounds symmetrically) both agree.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

FLT_EPSILON = 1.1920928955078125e-07          # <cfloat> FLT_EPSILON
FLT_MAX = 3.4028234663852886e38               # <cfloat> FLT_MAX
MAXLOC_THREADS = 512                          # cudaUtil.h:18 NTHREADS
SINC_PI = 3.14159265359                       # cuSincOverSampler.h:23-25

__all__ = [
    "FLT_EPSILON", "AmpcorConfig", "derived_sizes", "start_pixels", "check_range",
    "gather_chips", "amplitude", "subtract_mean", "correlate", "correlate_direct",
    "window_sums", "normalize", "maxloc2d", "estimate_variance", "estimate_snr",
    "adjust_offset", "deramp_method1", "oversample_c2c", "oversample_r2r",
    "sinc_filter_table", "oversample_sinc", "subpixel_offset", "run_windows",
    "run_ampcor",
]


# ---------------------------------------------------------------------------
# parameters (cuAmpcorParameter.cpp:18-109)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class AmpcorConfig:
    """The per-window knobs of ``cuAmpcorParameter`` (defaults: cuAmpcorParameter.cpp:18-64)."""
    window_h: int = 64            # windowSizeHeightRaw
    window_w: int = 64            # windowSizeWidthRaw
    hs_down: int = 20             # halfSearchRangeDownRaw
    hs_across: int = 20           # halfSearchRangeAcrossRaw
    raw_ovs: int = 2              # rawDataOversamplingFactor
    zoom: int = 16                # zoomWindowSize (before the setupParameters clamp)
    surf_ovs: int = 16            # oversamplingFactor
    surf_method: int = 0          # oversamplingMethod: 0 FFT, 1 sinc
    corr_stat: int = 21           # corrStatWindowSize
    deramp: int = 1               # derampMethod: 0 |z| chips, 1 deramp, other: complex, no deramp
    algorithm: int = 0            # 0 frequency domain, 1 time domain (same arithmetic here)


def derived_sizes(cfg: AmpcorConfig) -> dict:
    """Derived sizes of cuAmpcorParameter::setupParameters (cuAmpcorParameter.cpp:70-98)."""
    zoom = min(cfg.zoom, min(cfg.hs_across, cfg.hs_down) * 2 * cfg.raw_ovs)
    hz = zoom // (2 * cfg.raw_ovs)
    d = dict(
        corr_raw_zoom_h=min(cfg.corr_stat, 2 * cfg.hs_down + 1),
        corr_raw_zoom_w=min(cfg.corr_stat, 2 * cfg.hs_across + 1),
        zoom=zoom,
        hz=hz,
        os_win_h=cfg.window_h * cfg.raw_ovs,
        os_win_w=cfg.window_w * cfg.raw_ovs,
        search_h=cfg.window_h + 2 * cfg.hs_down,
        search_w=cfg.window_w + 2 * cfg.hs_across,
        zoom_in_h=cfg.window_h + 2 * hz,
        zoom_in_w=cfg.window_w + 2 * hz,
    )
    d["os_search_h"] = d["zoom_in_h"] * cfg.raw_ovs
    d["os_search_w"] = d["zoom_in_w"] * cfg.raw_ovs
    d["raw_corr_h"] = d["search_h"] - cfg.window_h + 1
    d["raw_corr_w"] = d["search_w"] - cfg.window_w + 1
    d["os_corr_h"] = d["os_search_h"] - d["os_win_h"] + 1
    d["os_corr_w"] = d["os_search_w"] - d["os_win_w"] + 1
    d["surf_in"] = zoom
    d["surf_out"] = zoom * cfg.surf_ovs
    if d["os_corr_h"] - 1 != zoom or d["os_corr_w"] - 1 != zoom:
        raise ValueError(
            f"corrSurfaceZoomInWindow {cfg.zoom} must be a multiple of 2*rawDataOversamplingFactor "
            f"(effective zoom {zoom}, trimmed surface {d['os_corr_h'] - 1})")
    return d


def start_pixels(n_down, n_across, start_down0, start_across0, skip_down, skip_across,
                 gross_down, gross_across, hs_down, hs_across):
    """Per-window reference/secondary start pixels (cuAmpcorParameter.cpp:170-205).

    ``gross_*`` are scalars (setConstantGrossOffset) or length-n arrays
    (setVaryingGrossOffset). Returns int64 arrays ``(refD, refA, secD, secA, gD, gA)``.
    """
    rows, cols = np.meshgrid(np.arange(n_down), np.arange(n_across), indexing="ij")
    ref_d = (start_down0 + rows * skip_down).ravel().astype(np.int64)
    ref_a = (start_across0 + cols * skip_across).ravel().astype(np.int64)
    g_d = np.broadcast_to(np.asarray(gross_down, np.int64), ref_d.shape).copy()
    g_a = np.broadcast_to(np.asarray(gross_across, np.int64), ref_a.shape).copy()
    return ref_d, ref_a, ref_d + g_d - hs_down, ref_a + g_a - hs_across, g_d, g_a


def check_range(ref_d, ref_a, sec_d, sec_a, cfg, ref_shape, sec_shape):
    """cuAmpcorParameter::checkPixelInImageRange (cuAmpcorParameter.cpp:275-331).

    Returns ``None`` when every window passes, else ``(window_index, message)``
    for the first failure. The end test is strict (``start + size >= dim`` fails),
    exactly as the reference.
    """
    d = derived_sizes(cfg)
    tests = [
        (ref_d < 0, "Reference Window start pixel out of range in Down"),
        (ref_a < 0, "Reference Window start pixel out of range in Across"),
        (ref_d + cfg.window_h >= ref_shape[0], "Reference Window end pixel out of range in Down"),
        (ref_a + cfg.window_w >= ref_shape[1], "Reference Window end pixel out of range in Across"),
        (sec_d < 0, "Secondary Window start pixel out of range in Down"),
        (sec_a < 0, "Secondary Window start pixel out of range in Across"),
        (sec_d + d["search_h"] >= sec_shape[0], "Secondary Window end pixel out of range in Down"),
        (sec_a + d["search_w"] >= sec_shape[1], "Secondary Window end pixel out of range in Across"),
    ]
    # the reference walks windows in order and stops at the first failing test of the first bad window
    bad = np.stack([np.asarray(b, bool) for b, _ in tests])
    any_bad = np.flatnonzero(bad.any(axis=0))
    if any_bad.size == 0:
        return None
    i = int(any_bad[0])
    return i, tests[int(np.argmax(bad[:, i]))][1]


# ---------------------------------------------------------------------------
# chip gather and elementwise stages
# ---------------------------------------------------------------------------
def gather_chips(image, starts_d, starts_a, h, w):
    """Gather ``[N, h, w]`` chips at per-window start pixels (cuArraysCopy.cu:70-82)."""
    image = np.asarray(image)
    ii = np.asarray(starts_d)[:, None, None] + np.arange(h)[None, :, None]
    jj = np.asarray(starts_a)[:, None, None] + np.arange(w)[None, None, :]
    return image[ii, jj]


def load_chips(image, starts_d, starts_a, h, w, deramp):
    """Chip load as in cuAmpcorChunk.cpp:280-317: complex rasters give complex chips, or
    ``(|z|, 0)`` when ``derampMethod == 0``; real rasters give ``(x, 0)``."""
    chips = gather_chips(image, starts_d, starts_a, h, w)
    if np.iscomplexobj(chips):
        chips = chips.astype(np.complex128)
        if deramp == 0:
            chips = np.abs(chips).astype(np.complex128)
        return chips
    return chips.astype(np.float64).astype(np.complex128)


def amplitude(chips):
    """|z| (cuArraysCopy.cu:510-529)."""
    return np.abs(chips)


def subtract_mean(images):
    """Per-image mean removal (cuCorrNormalization.cu:98-139)."""
    return images - images.mean(axis=(-2, -1), keepdims=True)


# ---------------------------------------------------------------------------
# correlation and normalization
# ---------------------------------------------------------------------------
def correlate(templates, searches):
    """Linear cross-correlation ``c[u,v] = sum_ij t[i,j] s[i+u,j+v]`` for
    ``u in [0, sh-th]``, ``v in [0, sw-tw]`` (cuCorrFrequency.cu:63-79: template
    zero-padded to the search size, ``conj(T)*S/|S|``, C2R, top-left crop;
    cuCorrTimeDomain.cu:12-92 computes the same sums directly)."""
    n, th, tw = templates.shape
    _, sh, sw = searches.shape
    tpad = np.zeros((n, sh, sw))
    tpad[:, :th, :tw] = templates
    spec = np.conj(np.fft.rfft2(tpad)) * np.fft.rfft2(searches)
    full = np.fft.irfft2(spec, s=(sh, sw))
    return full[:, : sh - th + 1, : sw - tw + 1]


def correlate_direct(templates, searches):
    """Brute-force version of :func:`correlate` (for small shapes)."""
    n, th, tw = templates.shape
    _, sh, sw = searches.shape
    out = np.zeros((n, sh - th + 1, sw - tw + 1))
    for u in range(sh - th + 1):
        for v in range(sw - tw + 1):
            out[:, u, v] = np.einsum("nij,nij->n", templates, searches[:, u:u + th, v:v + tw])
    return out


def window_sums(searches, th, tw):
    """Sums of ``s`` and ``s^2`` over every ``th x tw`` footprint, from FP64
    summed-area tables (cuCorrNormalizationSAT.cu:150-242)."""
    def box(a):
        sat = np.zeros((a.shape[0], a.shape[1] + 1, a.shape[2] + 1))
        sat[:, 1:, 1:] = a.cumsum(axis=1).cumsum(axis=2)
        return sat[:, th:, tw:] - sat[:, :-th, tw:] - sat[:, th:, :-tw] + sat[:, :-th, :-tw]
    return box(searches), box(searches * searches)


def normalize(corr, templates, searches):
    """NCC normalization (cuCorrNormalization.cu:230-314; same formula as the SAT
    path cuCorrNormalizationSAT.cu:237-240)."""
    th, tw = templates.shape[1:]
    t_sum2 = np.einsum("nij,nij->n", templates, templates)
    s1, s2 = window_sums(searches, th, tw)
    norm2 = (s2 - s1 * s1 / (th * tw)) * t_sum2[:, None, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        return corr / np.sqrt(norm2 + FLT_EPSILON)


# ---------------------------------------------------------------------------
# argmax, statistics, zoom adjustment, sub-pixel
# ---------------------------------------------------------------------------
def _bitrev(x, bits):
    r = np.zeros_like(x)
    for b in range(bits):
        r |= ((x >> b) & 1) << (bits - 1 - b)
    return r


def maxloc2d(surfaces):
    """Per-image argmax with the reference's tie rule (cuOffset.cu:23-84).

    Thread ``t`` of 512 scans ``idx = t, t+512, ...`` keeping the first strict
    maximum, then a strided tree keeps the left slot unless the right one is
    strictly greater. Among equal maxima this selects the smallest
    ``bitrev9(idx % 512)``, then the smallest ``idx``. Values not greater than
    ``-FLT_MAX`` (and NaN) are never selected; a surface with none yields
    index 0 and value ``-FLT_MAX`` (the reference leaves that case undefined).

    Returns ``(loc [N,2] int64 (row, col), val [N])``.
    """
    n, h, w = surfaces.shape
    flat = surfaces.reshape(n, h * w)
    ok = flat > -FLT_MAX
    masked = np.where(ok, flat, -np.inf)
    m = masked.max(axis=1)
    idx = np.arange(h * w, dtype=np.int64)
    key = _bitrev(idx % MAXLOC_THREADS, 9) * (h * w) + idx
    cand = (masked == m[:, None]) & ok
    keyc = np.where(cand, key[None, :], np.iinfo(np.int64).max)
    best = keyc.argmin(axis=1)
    none = ~ok.any(axis=1)
    best[none] = 0
    val = np.where(none, -FLT_MAX, m)
    return np.stack([best // w, best % w], axis=1), val


def estimate_variance(surf, loc, peak, template_size):
    """Curvature-based covariance (cuEstimateStats.cu:49-110). Returns ``[N,3]``
    ``(cov_down, cov_across, cov_cross)``; ``(99, 99, 0)`` at edges or when the
    Gaussian curvature ``|u| < 1e-2``."""
    n, nx, ny = surf.shape
    out = np.tile(np.array([99.0, 99.0, 0.0]), (n, 1))
    px, py = loc[:, 0], loc[:, 1]
    inner = (px - 1 >= 0) & (py - 1 >= 0) & (px + 1 < nx) & (py + 1 < ny)
    k = np.flatnonzero(inner)
    if k.size == 0:
        return out
    x, y, s = px[k], py[k], surf[k]
    g = lambda dx, dy: s[np.arange(k.size), x + dx, y + dy]  # noqa: E731
    dxx = -(g(1, 0) + g(-1, 0) - 2.0 * g(0, 0)) * template_size
    dyy = -(g(0, 1) + g(0, -1) - 2.0 * g(0, 0)) * template_size
    dxy = (g(1, 1) + g(-1, -1) - g(1, -1) - g(-1, 1)) * 0.25 * template_size
    n2 = np.maximum(1.0 - peak[k], 0.0)
    n4 = n2 * n2 * 0.5 * template_size
    n2 = n2 * 2.0
    u = dxy * dxy - dxx * dyy
    u2 = u * u
    good = np.abs(u) >= 1e-2
    with np.errstate(divide="ignore", invalid="ignore"):
        cxx = (-n2 * u * dyy + n4 * (dyy * dyy + dxy * dxy)) / u2
        cyy = (-n2 * u * dxx + n4 * (dxx * dxx + dxy * dxy)) / u2
        cxy = (n2 * u * dxy - n4 * (dxx + dyy) * dxy) / u2
    kk = k[good]
    out[kk, 0], out[kk, 1], out[kk, 2] = cxx[good], cyy[good], cxy[good]
    return out


def estimate_snr(surf, loc, peak, patch_h, patch_w):
    """SNR from the ``patch_h x patch_w`` patch centred on the peak
    (cuArraysCopy.cu:249-284 with ``-size/2`` centring and zero-fill,
    cuCorrNormalization.cu:143-192 sum of squares and valid count,
    cuEstimateStats.cu:20-46 ``max^2 / ((sum - max^2)/(count-1))``)."""
    n, nx, ny = surf.shape
    ii = loc[:, 0:1] + np.arange(patch_h)[None, :] - patch_h // 2
    jj = loc[:, 1:2] + np.arange(patch_w)[None, :] - patch_w // 2
    vi = (ii >= 0) & (ii < nx)
    vj = (jj >= 0) & (jj < ny)
    valid = vi[:, :, None] & vj[:, None, :]
    vals = surf[np.arange(n)[:, None, None], np.clip(ii, 0, nx - 1)[:, :, None],
                np.clip(jj, 0, ny - 1)[:, None, :]]
    vals = np.where(valid, vals, 0.0)
    s = (vals * vals).sum(axis=(1, 2))
    count = valid.sum(axis=(1, 2))
    with np.errstate(divide="ignore", invalid="ignore"):
        mean = (s - peak * peak) / (count - 1)
        return peak * peak / mean


def adjust_offset(maxloc, old_range, new_range):
    """Zoom-window start and residual shift per axis (cuOffset.cu:190-224)."""
    start = maxloc - new_range
    rbound = 2 * (old_range - new_range)
    shift = np.zeros_like(start)
    left = start < 0
    right = start > rbound
    shift = np.where(left, -start, np.where(right, start - rbound, 0))
    start = np.where(left, 0, np.where(right, rbound, start))
    return start, shift


def subpixel_offset(start, zoom_idx, surf_ovs, raw_ovs, hs_down, hs_across):
    """``offset = idx/(surf_ovs*raw_ovs) + start - hs`` per axis (cuOffset.cu:137-187)."""
    r = 1.0 / (surf_ovs * raw_ovs)
    return np.stack([r * zoom_idx[:, 0] + start[:, 0] - hs_down,
                     r * zoom_idx[:, 1] + start[:, 1] - hs_across], axis=1)


# ---------------------------------------------------------------------------
# deramp and FFT oversampling
# ---------------------------------------------------------------------------
def deramp_method1(chips):
    """Deramp method 1 (cuDeramp.cu:67-148) in FP64.

    ``phaseAcross = arg sum z[r,c] conj(z[r,c+1])``, ``phaseDown = arg sum
    z[r,c] conj(z[r+1,c])``; each pixel is multiplied by ``exp(i*phase)`` with
    ``phase = c*phaseDown + r*phaseAcross`` through float2.h's ``operator*=``,
    whose imaginary part reuses the updated real part.
    """
    z = np.asarray(chips, np.complex128)
    dy = (z[:, :, :-1] * np.conj(z[:, :, 1:])).sum(axis=(1, 2))
    dx = (z[:, :-1, :] * np.conj(z[:, 1:, :])).sum(axis=(1, 2))
    ph_y = np.arctan2(dy.imag, dy.real)
    ph_x = np.arctan2(dx.imag, dx.real)
    h, w = z.shape[1:]
    rows = np.arange(h)[None, :, None]
    cols = np.arange(w)[None, None, :]
    phase = cols * ph_x[:, None, None] + rows * ph_y[:, None, None]
    c, s = np.cos(phase), np.sin(phase)
    x, y = z.real, z.imag
    nx = x * c - y * s
    ny = y * c + nx * s
    return nx + 1j * ny


def _pad_quadrants(spec, out_h, out_w):
    """Quadrant padding (cuArraysPadding.cu:59-103): the ``h//2 x w//2`` corners
    of the spectrum move to the output corners, scaled by ``1/(h*w)``."""
    n, h, w = spec.shape
    h2, w2 = h // 2, w // 2
    f = 1.0 / (h * w)
    out = np.zeros((n, out_h, out_w), np.complex128)
    out[:, :h2, :w2] = spec[:, :h2, :w2] * f
    out[:, out_h - h2:, :w2] = spec[:, h - h2:, :w2] * f
    out[:, :h2, out_w - w2:] = spec[:, :h2, w - w2:] * f
    out[:, out_h - h2:, out_w - w2:] = spec[:, h - h2:, w - w2:] * f
    return out


def oversample_c2c(chips, out_h, out_w):
    """cuOverSamplerC2C::execute without the deramp (cuOverSampler.cpp:72-78):
    +i 2D DFT, quadrant pad, -i 2D DFT (both unnormalized)."""
    n, h, w = chips.shape
    spec = np.fft.ifft2(chips) * (h * w)            # +i sign, unnormalized
    return np.fft.fft2(_pad_quadrants(spec, out_h, out_w))


def oversample_r2r(surf, out_h, out_w):
    """cuOverSamplerR2R::execute (cuOverSampler.cpp:141-148) + real-part crop
    (cuArraysCopy.cu:326-339)."""
    return oversample_c2c(surf.astype(np.complex128), out_h, out_w).real


def sinc_filter_table():
    """Windowed-sinc table of cuSincOverSamplerR2R (cuSincOverSampler.cu:21-80,
    constants cuSincOverSampler.h:27-57): 8 taps x 4096 phases (+1)."""
    beta, relfiltlen, decfactor, pedestal = 0.75, 6.0, 4096, 0.0
    intplength = int(relfiltlen / beta + 0.5)
    filtercoef = intplength * decfactor
    wgthgt = (1.0 - pedestal) / 2.0
    soff = (filtercoef - 1.0) / 2.0
    i = np.arange(filtercoef + 1, dtype=np.float64)
    wa = i - soff
    wgt = (1.0 - wgthgt) + wgthgt * np.cos(SINC_PI * wa / soff)
    s = wa * beta / decfactor * SINC_PI
    with np.errstate(divide="ignore", invalid="ignore"):
        fct = np.where(s != 0.0, np.sin(s) / s, 1.0)
    return fct * wgt, intplength, decfactor


def oversample_sinc(surf, out_n, covs, center_shift, factor, window=2):
    """cuSincOverSamplerR2R::execute (cuSincOverSampler.cu:84-195): zero output
    except a ``(2*window*covs+1)^2`` patch around ``out/2 + shift*factor``."""
    n, in_x, in_y = surf.shape
    table, intplength, dec = sinc_filter_table()
    rng = window * covs
    start = out_n // 2 - rng
    size = 2 * rng + 1
    out = np.zeros((n, out_n, out_n))
    for k in range(n):
        ox = np.arange(size) + start + center_shift[k, 0] * factor
        oy = np.arange(size) + start + center_shift[k, 1] * factor
        ox = np.where(ox >= out_n, ox - out_n, ox)
        oy = np.where(oy >= out_n, oy - out_n, oy)

        def taps(o, n_in, n_wrap):
            io = o // covs
            frac = (o % covs) * (dec // covs) if dec % covs == 0 else \
                np.floor((o / covs - io) * dec).astype(np.int64)
            idx = io[:, None] - np.arange(intplength)[None, :] + intplength // 2
            idx = np.where(idx < 0, idx + n_in, idx)
            idx = np.where(idx >= n_in, idx - n_wrap, idx)
            coef = table[np.arange(intplength)[None, :] * dec + frac[:, None]]
            return idx, coef

        ix, cx = taps(ox, in_x, in_y)       # reference wraps x with inNY (cuSincOverSampler.cu:136)
        iy, cy = taps(oy, in_y, in_y)
        vals = surf[k][ix[:, None, :, None], iy[None, :, None, :]]     # [sx, sy, 8, 8]
        w = cx[:, None, :, None] * cy[None, :, None, :]
        out[k][np.ix_(ox, oy)] = (vals * w).sum(axis=(2, 3)) / w.sum(axis=(2, 3))
    return out


# ---------------------------------------------------------------------------
# the per-window pipeline (cuAmpcorChunk::run, cuAmpcorChunk.cpp:12-217)
# ---------------------------------------------------------------------------
def run_windows(ref_img, sec_img, ref_d, ref_a, sec_d, sec_a, cfg: AmpcorConfig,
                keep_stages=False, batch=64):
    """Run the pipeline for every window; returns a dict of per-window outputs.

    Keys: ``offset [N,2]`` (down, across), ``snr [N]``, ``cov [N,3]``,
    ``maxloc [N,2]`` (raw-surface argmax), ``maxval [N]``, ``start [N,2]``,
    ``shift [N,2]``, ``zoom_idx [N,2]`` (oversampled-surface argmax) and, with
    ``keep_stages``, the intermediate arrays named as the reference's
    ``CUAMPCOR_DEBUG`` dumps (cuAmpcorChunk.cpp:22-207).
    """
    d = derived_sizes(cfg)
    n = len(ref_d)
    outs = []
    for b0 in range(0, n, batch):
        sl = slice(b0, min(n, b0 + batch))
        outs.append(_run_batch(ref_img, sec_img, np.asarray(ref_d)[sl], np.asarray(ref_a)[sl],
                               np.asarray(sec_d)[sl], np.asarray(sec_a)[sl], cfg, d, keep_stages))
    return {k: np.concatenate([o[k] for o in outs]) for k in outs[0]}


def _run_batch(ref_img, sec_img, ref_d, ref_a, sec_d, sec_a, cfg, d, keep):
    st = {}
    W_h, W_w = cfg.window_h, cfg.window_w
    c_ref = load_chips(ref_img, ref_d, ref_a, W_h, W_w, cfg.deramp)
    r_ref = subtract_mean(amplitude(c_ref))
    c_sec = load_chips(sec_img, sec_d, sec_a, d["search_h"], d["search_w"], cfg.deramp)
    r_sec = amplitude(c_sec)
    corr = correlate(r_ref, r_sec)
    if keep:
        st.update(c_referenceBatchRaw=c_ref, r_referenceBatchRawSubMean=r_ref,
                  c_secondaryBatchRaw=c_sec, r_secondaryBatchRaw=r_sec, r_corrBatchRawUnNorm=corr)
    corr = normalize(corr, r_ref, r_sec)
    loc, peak = maxloc2d(corr)
    cov = estimate_variance(corr, loc, peak, W_h * W_w)
    snr = estimate_snr(corr, loc, peak, d["corr_raw_zoom_h"], d["corr_raw_zoom_w"])
    sx, shx = adjust_offset(loc[:, 0], cfg.hs_down, d["hz"])
    sy, shy = adjust_offset(loc[:, 1], cfg.hs_across, d["hz"])
    start = np.stack([sx, sy], axis=1)
    shift = np.stack([shx, shy], axis=1)
    if keep:
        st.update(r_corrBatchRaw=corr)

    ovs = cfg.raw_ovs
    c = deramp_method1(c_ref) if cfg.deramp == 1 else c_ref
    c_ref_os = oversample_c2c(c, d["os_win_h"], d["os_win_w"])
    r_ref_os = subtract_mean(amplitude(c_ref_os))
    zoom_chip = gather_chips_batch(c_sec, start, d["zoom_in_h"], d["zoom_in_w"])
    z = deramp_method1(zoom_chip) if cfg.deramp == 1 else zoom_chip
    c_sec_os = oversample_c2c(z, d["os_search_h"], d["os_search_w"])
    r_sec_os = amplitude(c_sec_os)
    corr_os = normalize(correlate(r_ref_os, r_sec_os), r_ref_os, r_sec_os)
    adj = corr_os[:, : d["zoom"], : d["zoom"]]
    if cfg.surf_method:
        surf = oversample_sinc(adj, d["surf_out"], cfg.surf_ovs, shift, cfg.surf_ovs * ovs)
    else:
        surf = oversample_r2r(adj, d["surf_out"], d["surf_out"])
    zidx, zval = maxloc2d(surf)
    off = subpixel_offset(start, zidx, cfg.surf_ovs, ovs, cfg.hs_down, cfg.hs_across)
    refined = None
    if not cfg.surf_method:
        refined = refine_offset(surf, zidx, start, cfg.surf_ovs, ovs, cfg.hs_down, cfg.hs_across)
    if keep:
        st.update(r_referenceBatchOverSampledSubMean=r_ref_os, c_secondaryBatchZoomIn=zoom_chip,
                  r_secondaryBatchOverSampled=r_sec_os, r_corrBatchZoomIn=corr_os,
                  r_corrBatchZoomInAdjust=adj, r_corrBatchZoomInOverSampled=surf)
    res = dict(offset=off, snr=snr, cov=cov, maxloc=loc, maxval=peak, start=start, shift=shift,
               zoom_idx=zidx, zoom_val=zval)
    if refined is not None:
        res["offset_fp64"] = refined
    res.update(st)
    return res


# ---------------------------------------------------------------------------
# FP64 sub-pixel refinement (no counterpart in the reference, whose offset is the
# bare argmax on the 1/(surf_ovs*raw_ovs) grid, cuOffset.cu:137-187)
# ---------------------------------------------------------------------------
def parabolic_vertex(fm, f0, fp):
    """Vertex of the parabola through (-1, fm), (0, f0), (1, fp), clamped to
    [-1/2, 1/2]; 0 where the three values do not bend down."""
    fm, f0, fp = (np.asarray(v, np.float64) for v in (fm, f0, fp))
    den = fm - 2.0 * f0 + fp
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(den < 0.0, 0.5 * (fm - fp) / den, 0.0)
    return np.clip(d, -0.5, 0.5)


def refine_offset(surf, zidx, start, surf_ovs, raw_ovs, hs_down, hs_across):
    """Offset refined by a parabola through the oversampled surface's peak and
    its two neighbours along each axis, in FP64: ``(idx + d)/(surf_ovs*raw_ovs)
    + start - hs``. The FFT-interpolated surface is periodic over its side, so
    the neighbours of an edge peak wrap around."""
    n, h, w = surf.shape
    k = np.arange(n)
    px, py = zidx[:, 0], zidx[:, 1]
    dx = parabolic_vertex(surf[k, (px - 1) % h, py], surf[k, px, py], surf[k, (px + 1) % h, py])
    dy = parabolic_vertex(surf[k, px, (py - 1) % w], surf[k, px, py], surf[k, px, (py + 1) % w])
    r = 1.0 / (surf_ovs * raw_ovs)
    return np.stack([(px + dx) * r + start[:, 0] - hs_down, (py + dy) * r + start[:, 1] - hs_across], axis=1)


def gather_chips_batch(chips, start, h, w):
    """Per-window crop at a varying start (cuArraysCopy.cu:209-224)."""
    n = chips.shape[0]
    ii = start[:, 0, None, None] + np.arange(h)[None, :, None]
    jj = start[:, 1, None, None] + np.arange(w)[None, None, :]
    return chips[np.arange(n)[:, None, None], ii, jj]


def run_ampcor(ref_img, sec_img, cfg: AmpcorConfig, n_down, n_across, start_down0, start_across0,
               skip_down, skip_across, gross_down=0, gross_across=0, merge_gross=False, **kw):
    """Whole-field run: start pixels, pipeline, and the controller's assembly
    (cuAmpcorController.cpp:130-158): offsets ``[nd, na, 2]`` (plus gross when
    ``merge_gross``), gross ``[nd, na, 2]``, snr ``[nd, na]``, cov ``[nd, na, 3]``."""
    ref_d, ref_a, sec_d, sec_a, g_d, g_a = start_pixels(
        n_down, n_across, start_down0, start_across0, skip_down, skip_across,
        gross_down, gross_across, cfg.hs_down, cfg.hs_across)
    r = run_windows(ref_img, sec_img, ref_d, ref_a, sec_d, sec_a, cfg, **kw)
    gross = np.stack([g_d, g_a], axis=1).astype(np.float64)
    off = r["offset"] + (gross if merge_gross else 0.0)
    r["offset_field"] = off.reshape(n_down, n_across, 2)
    r["gross_field"] = gross.reshape(n_down, n_across, 2)
    r["snr_field"] = r["snr"].reshape(n_down, n_across)
    r["cov_field"] = r["cov"].reshape(n_down, n_across, 3)
    return r
