"""Command-line dense offset estimation: the workflow around ``PyCuAmpcor``.

Same argument surface, window arithmetic, output names and sidecars as
contrib/PyCuAmpcor/examples/cuDenseOffsets.py, without isce/isceobj/GDAL:

    python -m isce2_amd.dense_offsets -r ref.slc -s sec.slc --op out/offset \
        --ww 64 --wh 64 --sw 20 --sh 20 --kw 32 --kh 32

* images are read through ``<name>.vrt`` / ``<name>.xml`` sidecars (or
  ``.npy``) with :mod:`isce2_amd.raster`; their size sets the window grid
  (cuDenseOffsets.py:181-238);
* ``-g 1 --gf FILE`` reads per-window int32 (down, across) gross offsets
  (cuDenseOffsets.py:299-321);
* outputs ``<prefix><suffix>.bip``, ``_gross.bip``, ``_snr.bip``, ``_cov.bip``
  with ISCE ``.xml`` and GDAL ``.vrt`` sidecars (cuDenseOffsets.py:274-281,
  353-395);
* ``--full-geom DIR --out-geom DIR`` resamples full-resolution geometry
  rasters onto the offset grid (prepareGeometry, cuDenseOffsets.py:400-460).

Differences: a gross-offset file of the wrong size raises ``ValueError``
(the reference prints a warning and returns 0, which its ``main`` then fails
to unpack); ``--fix-xml`` / ``--fix-vrt`` rewrite the sidecars with this
module's writers; ``--engine`` (auto/fused/staged) is an extension.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from pathlib import Path

import numpy as np

from . import raster

EXAMPLE = """example
  python -m isce2_amd.dense_offsets -r ./SLC/20151120/20151120.slc.full -s ./SLC/20151214/20151214.slc.full
  python -m isce2_amd.dense_offsets -r ref.slc -s sec.slc --outprefix ./offsets/offset --ww 256 --wh 256 \\
      --sw 8 --sh 8 --oo 32 --kw 300 --kh 100 --nwac 100 --nwdc 1 --gpuid 0
"""

GEOMETRY_BASES = ("lat", "lon", "los", "hgt", "z", "shadowMask", "waterMask")


def create_parser() -> argparse.ArgumentParser:
    """Argument surface of cuDenseOffsets.createParser (cuDenseOffsets.py:34-146)."""
    p = argparse.ArgumentParser(description="Dense pixel offsets between a reference and a secondary SLC "
                                            "on an MI355X GPU",
                                formatter_class=argparse.RawTextHelpFormatter, epilog=EXAMPLE)
    p.add_argument("-r", "--reference", dest="reference", required=True, help="reference SLC file")
    p.add_argument("-s", "--secondary", dest="secondary", required=True, help="secondary SLC file")
    p.add_argument("--fix-xml", "--fix-image-xml", dest="fixImageXml", action="store_true",
                   help="rewrite each input's .xml sidecar with absolute file paths")
    p.add_argument("--fix-vrt", "--fix-image-vrt", dest="fixImageVrt", action="store_true",
                   help="regenerate each input's .vrt sidecar from its .xml")
    p.add_argument("--op", "--outprefix", "--output-prefix", dest="outprefix", default="offset", required=True,
                   help="path prefix of every output file [%(default)s]")
    p.add_argument("--os", "--outsuffix", dest="outsuffix", default="", help="text appended to the prefix [%(default)s]")
    p.add_argument("--ww", type=int, dest="winwidth", default=64, help="columns per reference window [%(default)s]")
    p.add_argument("--wh", type=int, dest="winhgt", default=64, help="rows per reference window [%(default)s]")
    p.add_argument("--sw", type=int, dest="srcwidth", default=20,
                   help="search half-width across, in pixels [%(default)s]")
    p.add_argument("--sh", type=int, dest="srchgt", default=20,
                   help="search half-height down, in pixels [%(default)s]")
    p.add_argument("--kw", type=int, dest="skipwidth", default=64, help="column step between windows [%(default)s]")
    p.add_argument("--kh", type=int, dest="skiphgt", default=64, help="row step between windows [%(default)s]")
    p.add_argument("--mm", type=int, dest="margin", default=0, help="pixels left unused at each image edge [%(default)s]")
    p.add_argument("--nwa", type=int, dest="numWinAcross", default=-1,
                   help="window columns; -1 fits as many as the image allows [%(default)s]")
    p.add_argument("--nwd", type=int, dest="numWinDown", default=-1,
                   help="window rows; -1 fits as many as the image allows [%(default)s]")
    p.add_argument("--startpixelac", dest="startpixelac", type=int, default=-1,
                   help="column of the first reference window; -1 = margin plus search half-width [%(default)s]")
    p.add_argument("--startpixeldw", dest="startpixeldw", type=int, default=-1,
                   help="row of the first reference window; -1 = margin plus search half-height [%(default)s]")
    p.add_argument("--alg", "--algorithm", dest="algorithm", type=int, default=0,
                   help="correlation method: 0 spectral, 1 direct sums (same result)")
    p.add_argument("--raw-osf", "--raw-over-samp-factor", type=int, dest="raw_oversample", default=2,
                   choices=range(2, 5), help="chip interpolation factor before the fine search [%(default)s]")
    p.add_argument("--drmp", "--deramp", dest="deramp", type=int, default=0,
                   help="chip preparation: 0 amplitude, 1 remove the linear phase ramp [%(default)s]")
    g = p.add_argument_group("Gross offsets", "a constant or per-window prior shift of the search area")
    g.add_argument("-g", "--gross", type=int, dest="gross", default=0,
                   help="1: read per-window gross offsets from --gf [%(default)s]")
    g.add_argument("--aa", type=int, dest="azshift", default=0, help="constant gross shift down, pixels [%(default)s]")
    g.add_argument("--rr", type=int, dest="rgshift", default=0, help="constant gross shift across, pixels [%(default)s]")
    g.add_argument("--gf", "--gross-file", dest="gross_offset_file",
                   help="int32 (down, across) pairs, one per window, row-major")
    g.add_argument("--mg", "--merge-gross-offset", type=int, dest="merge_gross_offset", default=0,
                   help="1: write offsets with the gross shift added back [%(default)s]")
    c = p.add_argument_group("Correlation surface", "refinement of the coarse peak")
    c.add_argument("--corr-stat-size", type=int, dest="corr_stat_win_size", default=21,
                   help="side of the surface patch used for SNR and covariance [%(default)s]")
    c.add_argument("--corr-srch-size", type=int, dest="corr_srch_size", default=4,
                   help="half side, in raw pixels, of the surface patch that is refined [%(default)s]")
    c.add_argument("--corr-osf", "--oo", "--corr-over-samp-factor", type=int, dest="corr_oversample", default=32,
                   help="interpolation factor of the refined surface patch [%(default)s]")
    c.add_argument("--corr-osm", "--corr-over-samp-method", type=int, dest="corr_oversamplemethod", default=0,
                   help="surface interpolation: 0 spectral zero padding, 1 windowed sinc [%(default)s]")
    geo = p.add_argument_group("Geometry", "resample full-resolution geometry rasters onto the offset grid")
    geo.add_argument("--full-geom", dest="full_geometry_dir", help="directory holding the full-resolution *.rdr files")
    geo.add_argument("--out-geom", dest="out_geometry_dir", help="directory for the resampled geometry files")
    proc = p.add_argument_group("Execution")
    proc.add_argument("--gpuid", "--gid", "--gpu-id", dest="gpuid", type=int, default=0, help="HIP device index")
    proc.add_argument("--nstreams", dest="nstreams", type=int, default=2, help="accepted and ignored: the engine "
                      "schedules its own launches")
    proc.add_argument("--usemmap", dest="usemmap", type=int, default=1, help="read inputs through memory maps")
    proc.add_argument("--mmapsize", dest="mmapsize", type=int, default=8, help="accepted and ignored (GB)")
    proc.add_argument("--nwac", type=int, dest="numWinAcrossInChunk", default=10,
                      help="window columns per staged launch group [%(default)s]")
    proc.add_argument("--nwdc", type=int, dest="numWinDownInChunk", default=1,
                      help="window rows per staged launch group [%(default)s]")
    proc.add_argument("--redo", dest="redo", action="store_true", help="recompute when the offset file exists")
    proc.add_argument("--engine", dest="engine", default="auto", choices=["auto", "fused", "staged"],
                      help="kernel path (extension) [%(default)s]")
    return p


def parse_args(argv=None):
    return create_parser().parse_args(args=argv)


def image_size(name: str):
    """(length, width) of an image from its sidecar, as isceobj's SlcImage.load(name + '.xml')."""
    a = raster.open_raster(name)
    return int(a.shape[0]), int(a.shape[1])


def window_grid(length, width, inps):
    """Window counts and first reference pixel (cuDenseOffsets.py:215-238)."""
    margin = max(inps.margin, abs(inps.azshift), abs(inps.rgshift))
    nwd = inps.numWinDown if inps.numWinDown > 0 else (length - 2 * margin - 2 * inps.srchgt - inps.winhgt) // inps.skiphgt
    nwa = (inps.numWinAcross if inps.numWinAcross > 0
           else (width - 2 * margin - 2 * inps.srcwidth - inps.winwidth) // inps.skipwidth)
    sd = inps.startpixeldw if inps.startpixeldw != -1 else margin + inps.srchgt
    sa = inps.startpixelac if inps.startpixelac != -1 else margin + inps.srcwidth
    return dict(margin=margin, numberWindowDown=nwd, numberWindowAcross=nwa,
                referenceStartPixelDownStatic=sd, referenceStartPixelAcrossStatic=sa)


def read_gross_file(path: str, nwd: int, nwa: int):
    """Per-window int32 (down, across) gross offsets (cuDenseOffsets.py:303-318)."""
    g = np.fromfile(path, dtype=np.int32)
    n = nwd * nwa
    if g.size != 2 * n:
        raise ValueError(f"gross-offset file {path} holds {g.size} int32 values; a {nwd} x {nwa} window grid "
                         f"needs {2 * n}, so its size does not match the grid")
    g = g.reshape(n, 2)
    return np.ascontiguousarray(g[:, 0]), np.ascontiguousarray(g[:, 1])


def write_gross_file(path: str, down, across) -> None:
    """Inverse of :func:`read_gross_file` (int32 pairs, window order)."""
    np.stack([np.asarray(down, np.int32), np.asarray(across, np.int32)], axis=1).tofile(path)


def output_names(inps):
    fbase = f"{inps.outprefix}{inps.outsuffix}"
    return dict(offsetImageName=fbase + ".bip", grossOffsetImageName=fbase + "_gross.bip",
                snrImageName=fbase + "_snr.bip", covImageName=fbase + "_cov.bip")


def _render_sidecars(name, width, length, bands):
    raster.write_isce_xml(name, width, length, bands, "FLOAT", "BIP")
    raster.write_vrt(name, width, length, bands, "Float32", "BIP")


def _fix_sidecars(name, fix_xml, fix_vrt):
    """--fix-xml / --fix-vrt (cuDenseOffsets.py:166-179): re-render with absolute paths."""
    a = raster.open_raster(name)
    dt = "CFLOAT" if np.iscomplexobj(a) else "FLOAT"
    if fix_xml:
        raster.write_isce_xml(os.path.abspath(name), a.shape[1], a.shape[0], 1, dt, "BIP")
    if fix_vrt:
        raster.write_vrt(os.path.abspath(name), a.shape[1], a.shape[0], 1, "CFloat32" if dt == "CFLOAT" else "Float32")


def estimate_offset_field(reference: str, secondary: str, inps):
    """cuDenseOffsets.estimateOffsetField (cuDenseOffsets.py:156-397).

    Returns ``(obj, geom)``: the configured :class:`PyCuAmpcor` (after its run,
    unless the offsets exist and ``--redo`` is not given) and the offset grid's
    geometry (x/y start at the window centres, step, counts)."""
    from .PyCuAmpcor import PyCuAmpcor

    if inps.fixImageXml or inps.fixImageVrt:
        for fname in (reference, secondary):
            _fix_sidecars(fname, inps.fixImageXml, inps.fixImageVrt)
    length, width = image_size(reference)
    obj = PyCuAmpcor()
    obj.algorithm = inps.algorithm
    obj.deviceID = inps.gpuid
    obj.nStreams = inps.nstreams
    obj.derampMethod = inps.deramp
    obj.referenceImageName = _sidecar(reference)
    obj.referenceImageHeight, obj.referenceImageWidth = length, width
    obj.secondaryImageName = _sidecar(secondary)
    obj.secondaryImageHeight, obj.secondaryImageWidth = length, width
    grid = window_grid(length, width, inps)
    obj.numberWindowDown = grid["numberWindowDown"]
    obj.numberWindowAcross = grid["numberWindowAcross"]
    obj.windowSizeHeight, obj.windowSizeWidth = inps.winhgt, inps.winwidth
    obj.halfSearchRangeDown, obj.halfSearchRangeAcross = inps.srchgt, inps.srcwidth
    obj.referenceStartPixelDownStatic = grid["referenceStartPixelDownStatic"]
    obj.referenceStartPixelAcrossStatic = grid["referenceStartPixelAcrossStatic"]
    obj.skipSampleDown, obj.skipSampleAcross = inps.skiphgt, inps.skipwidth
    obj.rawDataOversamplingFactor = inps.raw_oversample
    obj.corrStatWindowSize = inps.corr_stat_win_size
    obj.corrSurfaceZoomInWindow = 2 * inps.corr_srch_size * inps.raw_oversample
    obj.corrSurfaceOverSamplingMethod = inps.corr_oversamplemethod
    obj.corrSurfaceOverSamplingFactor = inps.corr_oversample
    names = output_names(inps)
    for k, v in names.items():
        setattr(obj, k, v)
    obj.mergeGrossOffset = inps.merge_gross_offset
    obj.numberWindowDownInChunk = inps.numWinDownInChunk
    obj.numberWindowAcrossInChunk = inps.numWinAcrossInChunk
    obj.useMmap, obj.mmapSize = inps.usemmap, inps.mmapsize
    obj.engineMode = getattr(inps, "engine", "auto")
    obj.setupParams()
    if inps.gross == 0:
        obj.setConstantGrossOffset(inps.azshift, inps.rgshift)
    else:
        if not inps.gross_offset_file:
            raise ValueError("-g 1 needs --gross-file")
        gd, ga = read_gross_file(inps.gross_offset_file, obj.numberWindowDown, obj.numberWindowAcross)
        obj.setVaryingGrossOffset(gd, ga)
    obj.checkPixelInImageRange()
    geom = dict(x_start=obj.referenceStartPixelAcrossStatic + int(obj.windowSizeWidth / 2.0),
                y_start=obj.referenceStartPixelDownStatic + int(obj.windowSizeHeight / 2.0),
                x_step=obj.skipSampleAcross, y_step=obj.skipSampleDown,
                x_win_num=obj.numberWindowAcross, y_win_num=obj.numberWindowDown)
    if not inps.redo and os.path.exists(names["offsetImageName"]):
        print(f"{names['offsetImageName']} is already there; pass --redo to compute it again")
        return obj, geom
    obj.runAmpcor()
    nwa, nwd = obj.numberWindowAcross, obj.numberWindowDown
    _render_sidecars(names["offsetImageName"], nwa, nwd, 2)
    _render_sidecars(names["grossOffsetImageName"], nwa, nwd, 2)
    _render_sidecars(names["snrImageName"], nwa, nwd, 1)
    _render_sidecars(names["covImageName"], nwa, nwd, 3)
    return obj, geom


def _sidecar(name: str) -> str:
    """The raster the engine opens: ``name.vrt`` as the reference passes to GDAL
    (cuDenseOffsets.py:206,209), else the ``.xml`` descriptor, else the name."""
    for s in (name + ".vrt", name + ".xml"):
        if os.path.exists(s):
            return s
    return name


# ---------------------------------------------------------------------------
# geometry on the offset grid
# ---------------------------------------------------------------------------
_ENVI_TYPES = {np.dtype(np.uint8): 1, np.dtype(np.int16): 2, np.dtype(np.int32): 3, np.dtype(np.float32): 4,
               np.dtype(np.float64): 5, np.dtype(np.complex64): 6, np.dtype(np.complex128): 9}
_GDAL_NAMES = {np.dtype(np.uint8): "Byte", np.dtype(np.int16): "Int16", np.dtype(np.int32): "Int32",
               np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64",
               np.dtype(np.complex64): "CFloat32", np.dtype(np.complex128): "CFloat64"}


def nearest_subsample(band, x_start, y_start, x_step, y_step, x_num, y_num, nodata=0):
    """Nearest-neighbour resampling of the window [x_start, x_start + x_num*x_step) x
    [y_start, ...) onto x_num x y_num pixels, each taking the source pixel under its
    centre (start + floor((i + 0.5) * step)); pixels outside the band get ``nodata``
    (what gdal.Translate(srcWin, width, height, noData) does in prepareGeometry,
    cuDenseOffsets.py:444-450)."""
    h, w = band.shape
    xs = x_start + np.floor((np.arange(x_num) + 0.5) * x_step).astype(np.int64)
    ys = y_start + np.floor((np.arange(y_num) + 0.5) * y_step).astype(np.int64)
    out = np.full((y_num, x_num), nodata, dtype=band.dtype)
    vx, vy = (xs >= 0) & (xs < w), (ys >= 0) & (ys < h)
    out[np.ix_(vy, vx)] = np.asarray(band)[np.ix_(ys[vy], xs[vx])]
    return out


def _write_envi(path: str, bands):
    """Band-sequential raw file + ENVI header + VRT sidecar."""
    bands = [np.ascontiguousarray(b) for b in bands]
    h, w = bands[0].shape
    dt = bands[0].dtype
    with open(path, "wb") as f:
        for b in bands:
            b.tofile(f)
    with open(path + ".hdr", "w") as f:
        f.write(f"ENVI\ndescription = {{{os.path.basename(path)}}}\nsamples = {w}\nlines = {h}\nbands = {len(bands)}\n"
                f"header offset = 0\nfile type = ENVI Standard\ndata type = {_ENVI_TYPES[dt]}\n"
                f"interleave = bsq\nbyte order = 0\ndata ignore value = 0\n")
    raster.write_vrt(path, w, h, len(bands), _GDAL_NAMES[dt], "BSQ")


def prepare_geometry(full_dir, out_dir, x_start, y_start, x_step, y_step, x_win_num, y_win_num,
                     fbases=GEOMETRY_BASES):
    """prepareGeometry (cuDenseOffsets.py:400-460): geometry rasters on the offset grid.

    Reads ``<full_dir>/<base>.rdr.full`` (or ``.rdr`` when the directories
    differ) through its sidecar and writes ``<out_dir>/<base>.rdr`` (ENVI,
    band-sequential) with ``.hdr`` and ``.vrt``. Returns the written paths."""
    full_dir, out_dir = os.path.abspath(full_dir), os.path.abspath(out_dir)
    exts = [".rdr.full", ".rdr"] if full_dir != out_dir else [".rdr.full"]
    first = [e for e in exts if os.path.isfile(os.path.join(full_dir, fbases[0] + e))]
    if not first:
        raise ValueError(f"{full_dir} holds no {fbases[0]}.rdr.full or {fbases[0]}.rdr geometry raster")
    ext = first[0]
    os.makedirs(out_dir, exist_ok=True)
    written = []
    for base in fbases:
        src = os.path.join(full_dir, base + ext)
        if not os.path.isfile(src):
            continue
        bands = raster.open_bands(src)
        out = [nearest_subsample(b, x_start, y_start, x_step, y_step, x_win_num, y_win_num) for b in bands]
        dst = os.path.join(out_dir, base + ".rdr")
        _write_envi(dst, out)
        written.append(dst)
    return written


def main(argv=None):
    inps = parse_args(argv)
    t0 = time.time()
    out_dir = os.path.dirname(inps.outprefix)
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    geom = estimate_offset_field(inps.reference, inps.secondary, inps)[1]
    if inps.full_geometry_dir and inps.out_geometry_dir:
        prepare_geometry(inps.full_geometry_dir, inps.out_geometry_dir, **geom)
    m, s = divmod(time.time() - t0, 60)
    print(f"done in {int(m)} min {s:.1f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
