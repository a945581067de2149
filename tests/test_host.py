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
    with pytest.raises(IndexError, match=r"Reference Window end pixel out ot range in Across, window \(0,1\)"):
        p.checkPixelInImageRange()


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
