"""Host-side logic and the C ABI surface (no GPU needed)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from isce2_amd import PyCuAmpcor, raster
from isce2_amd import _lib
from oracle import ampcor_oracle as O

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    header = (ROOT / "include" / "isce2_amd.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(ampcor_\w+)\s*\(", header, re.M))
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.ampcor_version().decode().startswith("isce2_amd")


def test_create_fails_cleanly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    code = lib.ampcor_create(0, ctypes.byref(h))
    assert code != 0 and not h.value
    assert lib.ampcor_last_error(None)


def test_attribute_surface_and_defaults():
    p = PyCuAmpcor()
    # defaults of cuAmpcorParameter.cpp:18-64 through the PyCuAmpcor.cpp property names
    assert (p.algorithm, p.deviceID, p.nStreams, p.derampMethod) == (0, 0, 1, 1)
    assert (p.windowSizeHeight, p.windowSizeWidth, p.halfSearchRangeDown, p.halfSearchRangeAcross) == (64, 64, 20, 20)
    assert (p.skipSampleDown, p.skipSampleAcross, p.rawDataOversamplingFactor) == (64, 64, 2)
    assert (p.corrSurfaceZoomInWindow, p.corrSurfaceOverSamplingFactor, p.corrSurfaceOverSamplingMethod) == (16, 16, 0)
    assert (p.corrStatWindowSize, p.mergeGrossOffset, p.useMmap, p.mmapSize) == (21, 0, 1, 1)
    assert p.referenceImageName == "reference.slc" and p.offsetImageName == "DenseOffset.off"
    with pytest.raises(AttributeError):
        p.corrSufaceOverSamplingMethod = 1          # misspelling seen in a caller
    with pytest.raises(TypeError):
        p.windowSizeHeight = 64.0
    p.windowSizeHeight = np.int64(32)
    assert p.windowSizeHeight == 32 and isinstance(p.windowSizeHeight, int)
    p.grossOffsetAcrossStatic = 3                   # legacy alias accepted
    p.numberWindowDown, p.numberWindowAcross = 3, 4
    assert p.numberWindows == 12
    with pytest.raises(AttributeError):
        p.numberWindows = 5


def test_setup_and_start_pixels_match_oracle():
    p = PyCuAmpcor()
    p.numberWindowDown, p.numberWindowAcross = 3, 5
    p.halfSearchRangeDown = p.halfSearchRangeAcross = 3   # zoom clamp 16 -> 12
    p.referenceStartPixelDownStatic, p.referenceStartPixelAcrossStatic = 11, 7
    p.skipSampleDown, p.skipSampleAcross = 40, 30
    p.setupParams()
    assert p.corrSurfaceZoomInWindow == 12
    p.setConstantGrossOffset(2, -1)
    st = p._state
    rd, ra, sd, sa, gd, ga = O.start_pixels(3, 5, 11, 7, 40, 30, 2, -1, 3, 3)
    for k, v in zip(("ref_d", "ref_a", "sec_d", "sec_a", "gross_d", "gross_a"), (rd, ra, sd, sa, gd, ga)):
        np.testing.assert_array_equal(st[k], v)
    vd = np.arange(15, dtype=np.int32)
    va = -np.arange(15, dtype=np.int32)
    p.setVaryingGrossOffset(vd, va)
    np.testing.assert_array_equal(p._state["sec_d"], rd + vd - 3)
    np.testing.assert_array_equal(p._state["sec_a"], ra + va - 3)


def test_setup_errors_instead_of_exit():
    p = PyCuAmpcor()
    p.numberWindowDown = 0
    with pytest.raises(ValueError, match="Incorrect number of windows"):
        p.setupParams()


def test_check_range_strict_rule():
    p = PyCuAmpcor()
    p.referenceImageHeight = p.secondaryImageHeight = 200
    p.referenceImageWidth = p.secondaryImageWidth = 200
    p.numberWindowDown = p.numberWindowAcross = 2
    p.referenceStartPixelDownStatic = p.referenceStartPixelAcrossStatic = 20
    p.skipSampleDown = p.skipSampleAcross = 64
    p.setupParams()
    p.setConstantGrossOffset(0, 0)
    p.checkPixelInImageRange()                      # 20+64+104-20 = 168 < 200
    p.referenceImageWidth = 148                     # 84 + 64 = 148 >= 148 -> strict failure
    with pytest.raises(IndexError, match=r"Reference Window end pixel out of range in Across, window \(0,1\), "
                                         r"pixel 148"):
        p.checkPixelInImageRange()


def _c_params(wh, ww, hd, ha):
    return _lib.Params(window_h=wh, window_w=ww, hs_down=hd, hs_across=ha, raw_ovs=2, zoom=16, surf_ovs=16,
                       surf_method=0, corr_stat=21, deramp=1, algorithm=0, image_complex=1)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def test_c_abi_start_pixels_match_oracle():
    """ampcor_start_pixels (host memory, no device) against the oracle's
    setStartPixels restatement, for constant and varying gross offsets."""
    lib = _lib.load()
    rng = np.random.default_rng(7)
    for nd, na, s0d, s0a, kd, ka, hd, ha in [(3, 5, 11, 7, 40, 30, 3, 3), (7, 2, 0, 5, 17, 64, 20, 9),
                                              (1, 1, 100, 3, 1, 1, 16, 16)]:
        n = nd * na
        p = _c_params(64, 48, hd, ha)
        out = [np.full(n, -7, np.int32) for _ in range(4)]
        assert lib.ampcor_start_pixels(ctypes.byref(p), nd, na, s0d, s0a, kd, ka, 2, -1, None, None,
                                       *map(_p, out)) == 0
        ref = O.start_pixels(nd, na, s0d, s0a, kd, ka, 2, -1, hd, ha)
        for a, b in zip(out, ref[:4]):
            np.testing.assert_array_equal(a, b)
        gd = rng.integers(-30, 30, n).astype(np.int32)
        ga = rng.integers(-30, 30, n).astype(np.int32)
        assert lib.ampcor_start_pixels(ctypes.byref(p), nd, na, s0d, s0a, kd, ka, 0, 0, _p(gd), _p(ga),
                                       *map(_p, out)) == 0
        ref = O.start_pixels(nd, na, s0d, s0a, kd, ka, gd, ga, hd, ha)
        for a, b in zip(out, ref[:4]):
            np.testing.assert_array_equal(a, b)
    # argument errors come back as codes
    assert lib.ampcor_start_pixels(ctypes.byref(p), 0, 3, 0, 0, 1, 1, 0, 0, None, None, *map(_p, out)) == -1
    assert lib.ampcor_start_pixels(ctypes.byref(p), 1, 1, 0, 0, 1, 1, 0, 0, _p(gd), None, *map(_p, out)) == -1
    assert b"gross" in lib.ampcor_last_error(None)


def test_c_abi_check_range_every_test_at_its_edge():
    """ampcor_check_range against the oracle's strict rule: each of the eight
    tests, placed exactly at its boundary (passing) and one pixel beyond
    (failing), on a later window so that the reported index is checked too."""
    lib = _lib.load()
    wh, ww, hd, ha = 32, 40, 6, 9
    p = _c_params(wh, ww, hd, ha)
    cfg = O.AmpcorConfig(window_h=wh, window_w=ww, hs_down=hd, hs_across=ha)
    H, Wd = 300, 310
    base = [np.full(5, v, np.int32) for v in (50, 60, 44, 51)]
    # (array, value that just passes, value that just fails) for tests 0..7
    cases = [(0, 0, -1), (1, 0, -1), (0, H - wh - 1, H - wh), (1, Wd - ww - 1, Wd - ww),
             (2, 0, -1), (3, 0, -1), (2, H - (wh + 2 * hd) - 1, H - (wh + 2 * hd)),
             (3, Wd - (ww + 2 * ha) - 1, Wd - (ww + 2 * ha))]
    bad, which = ctypes.c_int32(), ctypes.c_int32()
    for k, (arr, ok_v, bad_v) in enumerate(cases):
        for v, expect in ((ok_v, 0), (bad_v, -2)):
            a = [b.copy() for b in base]
            a[arr][3] = v
            code = lib.ampcor_check_range(ctypes.byref(p), *map(_p, a), 5, H, Wd, H, Wd,
                                          ctypes.byref(bad), ctypes.byref(which))
            o = O.check_range(*[x.astype(np.int64) for x in a], cfg, (H, Wd), (H, Wd))
            assert code == expect, (k, v)
            if expect == 0:
                assert o is None and bad.value == -1
            else:
                assert (bad.value, which.value) == (3, k) and o[0] == 3
                assert lib.ampcor_last_error(None).decode().endswith("window 3")


def test_raster_readers_roundtrip(tmp_path):
    a = (np.arange(12 * 10).reshape(12, 10) + 1j * np.arange(120).reshape(12, 10)[::-1]).astype(np.complex64)
    fn = tmp_path / "img.slc"
    a.tofile(fn)
    raster.write_vrt(str(fn), 10, 12, 1, "CFloat32")
    raster.write_isce_xml(str(fn), 10, 12, 1, "CFLOAT")
    np.testing.assert_array_equal(raster.open_raster(str(fn) + ".vrt"), a)
    np.testing.assert_array_equal(raster.open_raster(str(fn) + ".xml"), a)
    np.testing.assert_array_equal(raster.open_raster(str(fn), 12, 10), a)
    # a 3-band BIP float file read through its VRT gives band 1
    b = np.random.default_rng(0).random((5, 7, 3)).astype(np.float32)
    fb = tmp_path / "cov.bip"
    b.tofile(fb)
    raster.write_vrt(str(fb), 7, 5, 3, "Float32", "BIP")
    np.testing.assert_array_equal(raster.open_raster(str(fb) + ".vrt"), b[:, :, 0])
    with pytest.raises(raster.RasterError):
        raster.open_raster(str(tmp_path / "missing.slc"))


def test_row_allocator_first_fit_and_coalescing():
    from isce2_amd.serving import RowAllocator
    a = RowAllocator(100)
    x = a.alloc(40)
    y = a.alloc(30)
    z = a.alloc(30)
    assert (x, y, z) == (0, 40, 70) and a.alloc(1) is None
    a.release(y, 30)
    assert a.alloc(31) is None and a.alloc(30) == 40
    a.release(x, 40)
    a.release(40, 30)
    assert a.free == [(0, 70)] and a.free_rows() == 70
    with pytest.raises(ValueError):
        a.release(10, 5)


def test_job_windows_grid_and_range_check():
    from isce2_amd.serving import OffsetJob, job_windows
    img = np.zeros((300, 260), np.complex64)
    j = OffsetJob(ref=img, sec=img, window=64, half_search=16, skip_down=32, skip_across=32)
    nd, na, rd, ra, sd, sa = job_windows(j)
    assert nd == (300 - 16 - 64 - 16) // 32 + 1 and na == (260 - 16 - 64 - 16) // 32 + 1
    assert rd[0] == 16 and ra[0] == 16 and sd[0] == 0 and sa[0] == 0
    assert (sd + 96).max() <= 300 and (sa + 96).max() <= 260
    j2 = OffsetJob(ref=img, sec=img, window=64, half_search=16, n_down=50, n_across=1)
    with pytest.raises(IndexError):
        job_windows(j2)


def test_fft_factorizations_multiply_to_length():
    from isce2_amd.autotune import factorizations
    for n in (16, 40, 48, 64, 72, 80, 96, 104, 136, 144, 160, 192, 256, 264, 272, 288, 320, 528):
        orders = factorizations(n)
        assert orders, n
        for r in orders:
            assert int(np.prod(r)) == n and set(r) <= {2, 3, 4, 5, 7, 8, 11, 13, 16, 17}
    assert factorizations(2 * 19) == []


def test_fill_rows_banded_copy_from_memmap(tmp_path):
    """The host staging copy of the banded uploader: any row range of a
    memory-mapped raster lands intact, on one thread or as row bands on a pool."""
    from concurrent.futures import ThreadPoolExecutor
    from isce2_amd.engine import fill_rows
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((3001, 1400)) + 1j * rng.standard_normal((3001, 1400))).astype(np.complex64)
    path = tmp_path / "img.c64"
    a.tofile(path)
    mm = np.memmap(path, dtype=np.complex64, mode="r", shape=a.shape)
    with ThreadPoolExecutor(5) as pool:
        for r0, r1, p, nthr in ((0, 3001, pool, 5), (17, 2990, pool, 5), (5, 9, pool, 5), (0, 3001, None, 1)):
            dst = np.zeros_like(a)
            fill_rows(dst, mm, r0, r1, p, nthr)
            assert np.array_equal(dst[r0:r1], a[r0:r1])
            assert not dst[:r0].any() and not dst[r1:].any()
