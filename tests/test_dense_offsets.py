"""Host side of the dense-offsets command line (cuDenseOffsets.py workflow).

Window-grid arithmetic on hand-worked cases, gross-offset files, geometry
resampling against brute-force indexing, sidecar round trips, and the
``--redo`` skip path, which runs without a GPU.
"""
import os

import numpy as np
import pytest

from isce2_amd import dense_offsets as D
from isce2_amd import raster


def _args(extra=()):
    return D.parse_args(["-r", "ref", "-s", "sec", "--op", "out/offset", *extra])


def test_parser_defaults_follow_the_reference():
    a = _args()
    # defaults of cuDenseOffsets.createParser (cuDenseOffsets.py:47-141)
    assert (a.winwidth, a.winhgt, a.srcwidth, a.srchgt, a.skipwidth, a.skiphgt) == (64, 64, 20, 20, 64, 64)
    assert (a.margin, a.numWinAcross, a.numWinDown, a.startpixelac, a.startpixeldw) == (0, -1, -1, -1, -1)
    assert (a.algorithm, a.raw_oversample, a.deramp, a.gross, a.azshift, a.rgshift) == (0, 2, 0, 0, 0, 0)
    assert (a.corr_stat_win_size, a.corr_srch_size, a.corr_oversample, a.corr_oversamplemethod) == (21, 4, 32, 0)
    assert (a.gpuid, a.nstreams, a.usemmap, a.mmapsize, a.numWinAcrossInChunk, a.numWinDownInChunk) == (0, 2, 1, 8, 10, 1)
    with pytest.raises(SystemExit):
        D.parse_args(["-r", "a", "-s", "b", "--op", "x", "--raw-osf", "5"])    # choices 2..4


def test_window_grid_hand_worked():
    # 1024 x 1024, window 64, search 20, skip 64: (1024 - 40 - 64) // 64 = 14 windows from pixel 20
    g = D.window_grid(1024, 1024, _args())
    assert (g["numberWindowDown"], g["numberWindowAcross"]) == (14, 14)
    assert (g["referenceStartPixelDownStatic"], g["referenceStartPixelAcrossStatic"]) == (20, 20)
    # a gross shift widens the margin: (1000 - 60 - 16 - 32) // 16 = 55 down, (700 - 60 - 40 - 32) // 24 = 23 across
    g = D.window_grid(1000, 700, _args(["--aa", "-30", "--rr", "7", "--ww", "32", "--wh", "32", "--sh", "8",
                                        "--kh", "16", "--kw", "24"]))
    assert (g["margin"], g["numberWindowDown"], g["numberWindowAcross"]) == (30, 55, 23)
    assert (g["referenceStartPixelDownStatic"], g["referenceStartPixelAcrossStatic"]) == (38, 50)
    g = D.window_grid(1000, 700, _args(["--nwd", "3", "--nwa", "4", "--startpixeldw", "5", "--startpixelac", "6"]))
    assert (g["numberWindowDown"], g["numberWindowAcross"]) == (3, 4)
    assert (g["referenceStartPixelDownStatic"], g["referenceStartPixelAcrossStatic"]) == (5, 6)


def test_gross_file_roundtrip_and_size_check(tmp_path):
    d = np.arange(12, dtype=np.int32) - 5
    a = -np.arange(12, dtype=np.int32)
    f = str(tmp_path / "gross.bin")
    D.write_gross_file(f, d, a)
    raw = np.fromfile(f, np.int32).reshape(12, 2)          # (down, across) per window
    assert np.array_equal(raw[:, 0], d) and np.array_equal(raw[:, 1], a)
    gd, ga = D.read_gross_file(f, 3, 4)
    assert np.array_equal(gd, d) and np.array_equal(ga, a)
    with pytest.raises(ValueError, match="does not match the grid"):
        D.read_gross_file(f, 4, 4)


def test_nearest_subsample_brute_force():
    rng = np.random.default_rng(0)
    band = rng.standard_normal((37, 53))
    for (x0, y0, xs, ys, nx, ny) in [(3, 5, 4, 3, 10, 8), (0, 0, 1, 1, 53, 37), (40, 30, 5, 4, 6, 5), (7, 2, 3, 7, 20, 9)]:
        out = D.nearest_subsample(band, x0, y0, xs, ys, nx, ny)
        ref = np.zeros((ny, nx))
        for i in range(ny):
            for j in range(nx):
                r, c = y0 + int((i + 0.5) * ys), x0 + int((j + 0.5) * xs)
                ref[i, j] = band[r, c] if r < 37 and c < 53 else 0.0
        assert np.array_equal(out, ref)


def test_prepare_geometry_multiband(tmp_path):
    full, out = tmp_path / "full", tmp_path / "geom"
    full.mkdir()
    rng = np.random.default_rng(1)
    h, w = 60, 80
    lat = rng.standard_normal((h, w))
    los = rng.standard_normal((2, h, w)).astype(np.float32)
    lat.tofile(full / "lat.rdr.full")
    raster.write_isce_xml(str(full / "lat.rdr.full"), w, h, 1, "DOUBLE", "BIP")
    np.stack([los[0], los[1]], axis=1).tofile(full / "los.rdr.full")          # BIL: [h][band][w]
    raster.write_isce_xml(str(full / "los.rdr.full"), w, h, 2, "FLOAT", "BIL")
    paths = D.prepare_geometry(str(full), str(out), x_start=4, y_start=6, x_step=8, y_step=5, x_win_num=9,
                               y_win_num=11)
    assert sorted(os.path.basename(p) for p in paths) == ["lat.rdr", "los.rdr"]
    rows = 6 + 5 * np.arange(11) + 2
    cols = 4 + 8 * np.arange(9) + 4
    got_lat = raster.open_bands(str(out / "lat.rdr"))
    assert len(got_lat) == 1 and np.array_equal(got_lat[0], lat[np.ix_(rows, cols)])
    got_los = raster.open_bands(str(out / "los.rdr"))
    assert len(got_los) == 2
    for b in range(2):
        assert np.array_equal(got_los[b], los[b][np.ix_(rows, cols)])
    hdr = (out / "los.rdr.hdr").read_text()
    assert "bands = 2" in hdr and "interleave = bsq" in hdr and "data type = 4" in hdr
    with pytest.raises(ValueError, match="geometry raster"):
        D.prepare_geometry(str(tmp_path / "empty"), str(out), 0, 0, 1, 1, 1, 1)


def test_existing_offsets_skip_the_run(tmp_path):
    # without --redo an existing offset file short-circuits before any device work
    h, w = 300, 400
    for name in ("ref.slc", "sec.slc"):
        (np.ones((h, w), np.complex64)).tofile(tmp_path / name)
        raster.write_vrt(str(tmp_path / name), w, h, 1, "CFloat32", "BIP")
    prefix = str(tmp_path / "o" / "offset")
    os.makedirs(tmp_path / "o")
    open(prefix + ".bip", "wb").close()
    inps = D.parse_args(["-r", str(tmp_path / "ref.slc"), "-s", str(tmp_path / "sec.slc"), "--op", prefix,
                         "--kw", "32", "--kh", "32"])
    obj, geom = D.estimate_offset_field(inps.reference, inps.secondary, inps)
    nwd, nwa = (300 - 40 - 64) // 32, (400 - 40 - 64) // 32
    assert (obj.numberWindowDown, obj.numberWindowAcross) == (nwd, nwa)
    assert geom == dict(x_start=20 + 32, y_start=20 + 32, x_step=32, y_step=32, x_win_num=nwa, y_win_num=nwd)
    assert obj.corrSurfaceZoomInWindow == 16 and obj.corrSurfaceOverSamplingFactor == 32
    assert obj.offsetImageName == prefix + ".bip" and obj.covImageName == prefix + "_cov.bip"
