"""Synthetic SLC scenes for offset tracking (speckle pairs with known shifts).

Generators for the workloads the engine is measured on:

* :func:`speckle_pair_shift` — circular complex Gaussian speckle smoothed by a
  3x3 box (circular, via FFT), secondary = reference shifted by a constant
  ``(down, across)`` through an FFT phase ramp, plus complex Gaussian noise at
  a given SNR. With 1024x1024, shift (+3.25, -1.75) and 10 dB this is the
  1024x1024 CPU-comparison scene.
* :func:`speckle_pair_warp` — the same speckle warped by a smooth offset field
  (2nd-order polynomial in row/col scaled to a maximum amplitude, plus
  ``N(0, jitter)`` per ``tile x tile`` block), resampled with a separable
  16-tap Kaiser-windowed sinc in FP64 (first along down, then across). With
  8192x4096 and +-8 px this is the 8192x4096 plain-GPU scene.
* :func:`chip_pairs_uneven` — independent chip pairs of mixed window size and
  search range with sub-pixel shifts (the uneven-batch workload).

Sign convention: a feature at ``(r, c)`` in the reference appears at
``(r + d_down, c + d_across)`` in the secondary, so the offset tracker should
return ``+d``. All generators are seed-fixed and run on CPU or on a GPU via
torch (FP64 arithmetic, complex64 output).
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _dev(device):
    if device is None:
        return torch.device("cpu")
    return torch.device(device)


def speckle(h, w, seed, device=None):
    """Complex Gaussian speckle (re, im ~ N(0, 1/2)) smoothed by a circular 3x3 box, complex128."""
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    re = torch.randn(h, w, generator=g, dtype=torch.float64)
    im = torch.randn(h, w, generator=g, dtype=torch.float64)
    z = torch.complex(re, im).to(_dev(device)) * math.sqrt(0.5)
    # circular 3x3 box via FFT: transfer function of the box is separable
    fy = torch.fft.fftfreq(h, device=z.device, dtype=torch.float64)
    fx = torch.fft.fftfreq(w, device=z.device, dtype=torch.float64)
    hy = (1 + 2 * torch.cos(2 * math.pi * fy)) / 3.0
    hx = (1 + 2 * torch.cos(2 * math.pi * fx)) / 3.0
    return torch.fft.ifft2(torch.fft.fft2(z) * (hy[:, None] * hx[None, :]))


def _noise(shape, power, seed, device):
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    re = torch.randn(*shape, generator=g, dtype=torch.float64)
    im = torch.randn(*shape, generator=g, dtype=torch.float64)
    return (torch.complex(re, im) * math.sqrt(power / 2.0)).to(_dev(device))


def fft_shift(z, d_down, d_across):
    """Shift a complex image by a constant (sub-pixel) amount with an FFT phase ramp."""
    h, w = z.shape[-2:]
    fy = torch.fft.fftfreq(h, device=z.device, dtype=torch.float64)
    fx = torch.fft.fftfreq(w, device=z.device, dtype=torch.float64)
    ramp = torch.exp(-2j * math.pi * (fy[:, None] * d_down + fx[None, :] * d_across))
    return torch.fft.ifft2(torch.fft.fft2(z) * ramp)


def speckle_pair_shift(h=1024, w=1024, shift=(3.25, -1.75), snr_db=10.0, seed=0, device=None):
    """Reference/secondary complex64 numpy arrays with a constant shift and additive noise."""
    ref = speckle(h, w, seed, device)
    sec = fft_shift(ref, shift[0], shift[1])
    p = float((ref.abs() ** 2).mean())
    if snr_db is not None:
        sec = sec + _noise((h, w), p / 10 ** (snr_db / 10.0), seed + 1, ref.device)
    return (ref.to(torch.complex64).cpu().numpy(), sec.to(torch.complex64).cpu().numpy())


def _kaiser_sinc(frac, taps=16, beta=6.0):
    """Separable Kaiser-windowed sinc weights for fractional positions ``frac`` in [0,1)."""
    k = torch.arange(taps, device=frac.device, dtype=torch.float64) - (taps // 2 - 1)   # -7..8
    x = k[None, :] - frac[..., None]
    win_arg = 1 - (x / (taps / 2)) ** 2
    win = torch.special.i0(beta * torch.sqrt(torch.clamp(win_arg, min=0.0))) / torch.special.i0(
        torch.tensor(beta, dtype=torch.float64, device=frac.device))
    return torch.sinc(x) * win, k


def _resample_axis(z, disp, axis, taps=16, block=256):
    """out[r] = z(r - disp[r]) along ``axis`` with a windowed sinc (zero outside)."""
    if axis == 1:
        return _resample_axis(z.transpose(0, 1), disp.transpose(0, 1), 0, taps, block).transpose(0, 1)
    h, w = z.shape
    out = torch.empty_like(z)
    cols = torch.arange(w, device=z.device)
    for r0 in range(0, h, block):
        r1 = min(h, r0 + block)
        pos = torch.arange(r0, r1, device=z.device, dtype=torch.float64)[:, None] - disp[r0:r1]
        base = torch.floor(pos)
        wts, k = _kaiser_sinc(pos - base, taps)
        idx = base.long()[..., None] + k.long()[None, None, :]
        valid = (idx >= 0) & (idx < h)
        vals = z[idx.clamp(0, h - 1), cols[None, :, None].expand_as(idx)]
        out[r0:r1] = (vals * torch.where(valid, wts, torch.zeros_like(wts))).sum(-1)
    return out


def offset_field(h, w, max_amp=8.0, jitter=0.05, tile=256, seed=2, device=None):
    """Smooth offset field: 2nd-order polynomial in (row/H, col/W) scaled to max |d| = max_amp,
    plus N(0, jitter) per tile. Returns float64 tensors (d_down, d_across) of shape (h, w)."""
    dev = _dev(device)
    rng = np.random.default_rng(seed)
    y = torch.arange(h, device=dev, dtype=torch.float64)[:, None] / h
    x = torch.arange(w, device=dev, dtype=torch.float64)[None, :] / w
    basis = [torch.ones_like(y * x), y + 0 * x, x + 0 * y, y * y + 0 * x, x * y, x * x + 0 * y]
    out = []
    for comp in range(2):
        c = rng.uniform(-1, 1, size=6)
        d = sum(float(ci) * b for ci, b in zip(c, basis))
        d = d * (max_amp / float(d.abs().max()))
        jt = torch.as_tensor(rng.normal(0, jitter, size=((h + tile - 1) // tile, (w + tile - 1) // tile)),
                             device=dev)
        d = d + jt.repeat_interleave(tile, 0)[:h].repeat_interleave(tile, 1)[:, :w]
        out.append(d)
    return out[0], out[1]


def speckle_pair_warp(h=8192, w=4096, max_amp=8.0, jitter=0.05, tile=256, seed=0, device=None):
    """Reference/secondary complex64 numpy arrays, secondary warped by a smooth offset field.

    Returns ``(ref, sec, d_down, d_across)`` (offset fields as float32 numpy)."""
    ref = speckle(h, w, seed, device)
    dd, da = offset_field(h, w, max_amp, jitter, tile, seed + 2, device)
    sec = _resample_axis(_resample_axis(ref, dd, 0), da, 1)
    return (ref.to(torch.complex64).cpu().numpy(), sec.to(torch.complex64).cpu().numpy(),
            dd.float().cpu().numpy(), da.float().cpu().numpy())


UNEVEN_WINDOWS = (32, 64, 128, 256)
UNEVEN_RANGES = (8, 16, 32)


def chip_pairs_uneven(n=8192, seed=0, windows=UNEVEN_WINDOWS, ranges=UNEVEN_RANGES, device=None):
    """Independent chip pairs of mixed sizes.

    Window ``W`` uniform over ``windows``, half search ``hs`` uniform over
    ``ranges``, true offsets uniform in +-0.8*hs per axis. Each pair is cut from
    its own speckle field shifted by an FFT phase ramp. Returns a list of dicts
    ``{W, hs, ref [W,W] c64, sec [W+2hs, W+2hs] c64, shift (down, across)}``.
    """
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        W = int(rng.choice(windows))
        hs = int(rng.choice(ranges))
        sh = rng.uniform(-0.8 * hs, 0.8 * hs, size=2)
        S = W + 2 * hs
        field = speckle(S + 32, S + 32, seed * 1_000_003 + i, device)
        moved = fft_shift(field, sh[0], sh[1])
        ref = field[16 + hs:16 + hs + W, 16 + hs:16 + hs + W]
        sec = moved[16:16 + S, 16:16 + S]
        pairs.append(dict(W=W, hs=hs, shift=(float(sh[0]), float(sh[1])),
                          ref=ref.to(torch.complex64).cpu().numpy(),
                          sec=sec.to(torch.complex64).cpu().numpy()))
    return pairs


def chip_pairs_uneven_grouped(n=8192, seed=0, windows=UNEVEN_WINDOWS, ranges=UNEVEN_RANGES, device=None,
                              block=128):
    """The uneven-batch workload generated per size class, vectorized.

    Window ``W`` uniform over ``windows``, half search ``hs`` uniform over
    ``ranges``, true offsets uniform in +-0.8*hs per axis (continuous). Each
    pair comes from its own speckle field (smoothed by the 3x3 box) and its
    secondary is that field shifted by an FFT phase ramp. Returns
    ``{(W, hs): {"idx": int64[m], "ref": complex64[m, W, W], "sec":
    complex64[m, S, S], "shift": float64[m, 2]}}`` with tensors on ``device``.
    """
    dev = _dev(device)
    rng = np.random.default_rng(seed)
    Ws = rng.choice(np.asarray(windows), size=n)
    hss = rng.choice(np.asarray(ranges), size=n)
    shifts = rng.uniform(-0.8, 0.8, size=(n, 2)) * hss[:, None]
    out = {}
    for W in windows:
        for hs in ranges:
            idx = np.nonzero((Ws == W) & (hss == hs))[0]
            if len(idx) == 0:
                continue
            S = W + 2 * hs
            F = S + 32
            g = torch.Generator(device="cpu").manual_seed(int(seed) * 7919 + int(W) * 131 + int(hs))
            fy = torch.fft.fftfreq(F, device=dev, dtype=torch.float64)
            box = (1 + 2 * torch.cos(2 * math.pi * fy)) / 3.0
            refs, secs = [], []
            for b0 in range(0, len(idx), block):
                m = min(block, len(idx) - b0)
                re = torch.randn(m, F, F, generator=g, dtype=torch.float64)
                im = torch.randn(m, F, F, generator=g, dtype=torch.float64)
                z = torch.complex(re, im).to(dev) * math.sqrt(0.5)
                spec = torch.fft.fft2(z) * (box[:, None] * box[None, :])
                field = torch.fft.ifft2(spec)
                sh = torch.as_tensor(shifts[idx[b0:b0 + m]], device=dev)
                ramp = torch.exp(-2j * math.pi * (fy[None, :, None] * sh[:, 0, None, None] +
                                                  fy[None, None, :] * sh[:, 1, None, None]))
                moved = torch.fft.ifft2(spec * ramp)
                refs.append(field[:, 16 + hs:16 + hs + W, 16 + hs:16 + hs + W].to(torch.complex64))
                secs.append(moved[:, 16:16 + S, 16:16 + S].to(torch.complex64))
            out[(int(W), int(hs))] = dict(idx=idx, ref=torch.cat(refs), sec=torch.cat(secs), shift=shifts[idx])
    return out


def precision_stress_scene(size=4096, tile=128, sigma=2.0, n_points=200, point_db=60.0, tiny_frac=0.2,
                           tiny_amp=1e-6, max_shift=4.0, seed=0, device=None):
    """Precision-stress pair: heavy-tailed amplitude, bright points, near-zero windows.

    Reference = unit speckle (3x3 box) times a per-pixel log-normal amplitude
    ``exp(sigma * N(0, 1))``, plus ``n_points`` band-limited point targets of
    amplitude ``10**(point_db/20)`` x the mean amplitude at random sub-pixel
    positions. The image is cut into ``tile x tile`` tiles; each tile's
    secondary is the tile shifted by its own ``U(+-max_shift)`` offset through
    an FFT phase ramp (circular within the tile), and a ``tiny_frac`` share of
    tiles (reference and secondary) is scaled to ``tiny_amp``. Returns
    ``(ref, sec, shifts [nt, nt, 2], tiny [nt, nt] bool)`` as numpy arrays.
    """
    dev = _dev(device)
    rng = np.random.default_rng(seed)
    nt = size // tile
    z = speckle(size, size, seed, dev)
    g = torch.Generator(device="cpu").manual_seed(int(seed) + 17)
    amp = torch.exp(sigma * torch.randn(size, size, generator=g, dtype=torch.float64)).to(dev)
    z = z * amp
    mean_amp = float(z.abs().mean())
    t = z.reshape(nt, tile, nt, tile).permute(0, 2, 1, 3)                  # [nt, nt, tile, tile]
    spec = torch.fft.fft2(t)
    f = torch.fft.fftfreq(tile, device=dev, dtype=torch.float64)
    py = rng.uniform(0, size, n_points)
    px = rng.uniform(0, size, n_points)
    a_pt = mean_amp * 10 ** (point_db / 20.0)
    for y, x in zip(py, px):
        ty, tx = int(y // tile), int(x // tile)
        yl, xl = y - ty * tile, x - tx * tile
        # band-limited impulse of peak amplitude a_pt at (yl, xl): spectrum a_pt * e^{-2 pi i f.r}
        spec[ty, tx] += a_pt * torch.exp(-2j * math.pi * (f[:, None] * yl + f[None, :] * xl))
    shifts = rng.uniform(-max_shift, max_shift, size=(nt, nt, 2))
    sh = torch.as_tensor(shifts, device=dev)
    ramp = torch.exp(-2j * math.pi * (f[None, None, :, None] * sh[..., 0, None, None] +
                                      f[None, None, None, :] * sh[..., 1, None, None]))
    ref_t = torch.fft.ifft2(spec)
    sec_t = torch.fft.ifft2(spec * ramp)
    tiny = rng.uniform(size=(nt, nt)) < tiny_frac
    scale = torch.where(torch.as_tensor(tiny, device=dev), tiny_amp, 1.0)[..., None, None]
    ref_t = ref_t * scale
    sec_t = sec_t * scale
    back = lambda u: u.permute(0, 2, 1, 3).reshape(size, size).to(torch.complex64).cpu().numpy()  # noqa: E731
    return back(ref_t), back(sec_t), shifts, tiny


def window_grid(h, w, window, hs, skip, margin=0):
    """Window counts and first start pixel as cuDenseOffsets computes them
    (contrib/PyCuAmpcor/examples/cuDenseOffsets.py:221-241)."""
    nd = (h - 2 * margin - 2 * hs - window) // skip
    na = (w - 2 * margin - 2 * hs - window) // skip
    return nd, na, margin + hs, margin + hs
