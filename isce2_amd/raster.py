"""Raster I/O without GDAL: the inputs and outputs of PyCuAmpcor's runAmpcor.

Readers (all memory-mapped, single band, row-major):

* GDAL ``.vrt`` with one ``VRTRawRasterBand`` (what ISCE writes next to its
  raw binaries): ``SourceFilename`` (``relativeToVRT``), ``ImageOffset``,
  ``PixelOffset``, ``LineOffset``, ``ByteOrder``; data types ``CFloat32``
  and ``Float32``. Replaces GDALImage's open + band lookup
  (contrib/PyCuAmpcor/src/GDALImage.cpp:23-94).
* ISCE ``.xml`` image descriptors (``width``, ``length``, ``data_type``,
  ``file_name``; ``scheme`` BIP/BIL/BSQ with one band).
* ``.npy`` arrays.
* raw binaries given explicit height/width (the element size decides between
  complex64 and float32).

Writers produce the reference's raw outputs (native-endian float32, BIP;
cuArrays.cpp:162-169) and optionally ISCE ``.xml`` / GDAL ``.vrt`` sidecars
as cuDenseOffsets.py:353-395 renders them.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

_GDAL_TYPES = {"CFloat32": np.complex64, "Float32": np.float32, "CFloat64": np.complex128,
               "Float64": np.float64, "Int16": np.int16, "CInt16": None, "Byte": np.uint8}
_ISCE_TYPES = {"CFLOAT": np.complex64, "FLOAT": np.float32, "DOUBLE": np.float64,
               "CDOUBLE": np.complex128, "SHORT": np.int16, "BYTE": np.uint8}


class RasterError(IOError):
    pass


def _vrt_band(path: Path, b, w: int, h: int):
    if b.get("subClass") != "VRTRawRasterBand":
        raise RasterError(f"{path}: only VRTRawRasterBand is supported")
    dt = _GDAL_TYPES.get(b.get("dataType"))
    if dt is None:
        raise RasterError(f"{path}: unsupported dataType {b.get('dataType')}")
    src = b.find("SourceFilename")
    fn = src.text.strip()
    if src.get("relativeToVRT", "0") == "1" or not os.path.isabs(fn):
        fn = str(path.parent / fn)
    item = np.dtype(dt).itemsize
    img_off = int((b.findtext("ImageOffset") or "0").strip())
    pix_off = int((b.findtext("PixelOffset") or str(item)).strip())
    line_off = int((b.findtext("LineOffset") or str(item * w)).strip())
    order = (b.findtext("ByteOrder") or "LSB").strip().upper()
    dtype = np.dtype(dt).newbyteorder("<" if order == "LSB" else ">")
    if pix_off % item or line_off % item:
        raise RasterError(f"{path}: pixel/line offsets not multiples of the element size")
    raw = np.memmap(fn, dtype=dtype, mode="r", offset=img_off)
    strides = (line_off, pix_off)
    need = (h - 1) * line_off + (w - 1) * pix_off + item
    if raw.size * item < need:
        raise RasterError(f"{fn}: file smaller than the VRT geometry")
    return np.lib.stride_tricks.as_strided(raw, shape=(h, w), strides=strides, writeable=False)


def _vrt_bands(path: Path):
    root = ET.parse(path).getroot()
    w, h = int(root.get("rasterXSize")), int(root.get("rasterYSize"))
    bands = root.findall("VRTRasterBand")
    if not bands:
        raise RasterError(f"{path}: no VRTRasterBand")
    return [_vrt_band(path, b, w, h) for b in bands]


def _from_vrt(path: Path):
    return _vrt_bands(path)[0]


def _isce_xml_bands(path: Path):
    root = ET.parse(path).getroot()
    props = {}
    for p in root.iter("property"):
        name = (p.get("name") or "").lower()
        val = p.findtext("value")
        if val is not None:
            props[name] = val.strip()
    w, h = int(props["width"]), int(props["length"])
    dt = _ISCE_TYPES.get(props.get("data_type", "CFLOAT").upper())
    if dt is None:
        raise RasterError(f"{path}: unsupported data_type")
    nb = int(props.get("number_bands", "1"))
    fn = props.get("file_name", str(path)[:-4])
    if not os.path.isabs(fn) or not os.path.exists(fn):
        fn = str(path.parent / os.path.basename(fn))
    order = props.get("byte_order", "l").lower()
    dtype = np.dtype(dt).newbyteorder("<" if order.startswith("l") else ">")
    scheme = props.get("scheme", "BIP").upper()
    if scheme == "BIL":
        a = np.memmap(fn, dtype=dtype, mode="r", shape=(h, nb, w))
        return [a[:, i, :] for i in range(nb)]
    if scheme == "BSQ":
        a = np.memmap(fn, dtype=dtype, mode="r", shape=(nb, h, w))
        return [a[i] for i in range(nb)]
    a = np.memmap(fn, dtype=dtype, mode="r", shape=(h, w, nb))
    return [a[:, :, i] for i in range(nb)]


def _from_isce_xml(path: Path):
    bands = _isce_xml_bands(path)
    if len(bands) != 1:
        raise RasterError(f"{path}: only single-band images are supported")
    return bands[0]


def open_bands(name: str):
    """Every band of a multi-band raster (``.vrt`` or ISCE ``.xml`` sidecar, BIP/BIL/BSQ)
    as read-only 2-D arrays; single-band formats give a one-element list."""
    p = Path(name)
    if p.suffix == ".vrt" and p.exists():
        return _vrt_bands(p)
    if p.suffix == ".xml" and p.exists():
        return _isce_xml_bands(p)
    for side in (Path(str(p) + ".vrt"), Path(str(p) + ".xml")):
        if side.exists():
            return open_bands(str(side))
    return [open_raster(name)]


def open_raster(name: str, height: int | None = None, width: int | None = None):
    """Open ``name`` as a read-only 2-D array (memory-mapped where possible)."""
    p = Path(name)
    if p.suffix == ".vrt" and p.exists():
        return _from_vrt(p)
    if p.suffix == ".xml" and p.exists():
        return _from_isce_xml(p)
    if p.suffix == ".npy" and p.exists():
        a = np.load(p, mmap_mode="r", allow_pickle=False)
        if a.ndim != 2:
            raise RasterError(f"{p}: expected a 2-D array")
        return a
    for side in (Path(str(p) + ".vrt"), Path(str(p) + ".xml")):
        if side.exists():
            return open_raster(str(side))
    if not p.exists():
        raise RasterError(f"cannot open raster {name}")
    if not height or not width:
        raise RasterError(f"{name}: raw binary needs image height and width")
    size = p.stat().st_size
    per = size // (int(height) * int(width))
    if per == 8:
        dt = np.complex64
    elif per == 4:
        dt = np.float32
    else:
        raise RasterError(f"{name}: {size} bytes is not {height}x{width} complex64 or float32")
    return np.memmap(p, dtype=dt, mode="r", shape=(int(height), int(width)))


def write_raw(name: str, array) -> None:
    """Write native-endian raw bytes (cuArrays.cpp:162-169)."""
    np.ascontiguousarray(array, dtype=np.float32).tofile(name)


def write_isce_xml(name: str, width: int, length: int, bands: int, data_type: str = "FLOAT",
                   scheme: str = "BIP") -> None:
    """Minimal ISCE image descriptor (``<file>.xml``) as isceobj Image.renderHdr writes."""
    props = dict(file_name=os.path.abspath(name), width=width, length=length, number_bands=bands,
                 data_type=data_type, scheme=scheme, byte_order="l", access_mode="read",
                 image_type="image")
    root = ET.Element("imageFile")
    for k, v in props.items():
        pe = ET.SubElement(root, "property", name=k)
        ET.SubElement(pe, "value").text = str(v)
    ET.ElementTree(root).write(name + ".xml")


def write_vrt(name: str, width: int, length: int, bands: int, data_type: str = "Float32",
              scheme: str = "BIP") -> None:
    """GDAL VRTRawRasterBand sidecar (isceobj/Image/Image.py:461-579 layout rules)."""
    item = {"Byte": 1, "Int16": 2, "Int32": 4, "Float32": 4, "Float64": 8, "CFloat32": 8, "CFloat64": 16}[data_type]
    root = ET.Element("VRTDataset", rasterXSize=str(width), rasterYSize=str(length))
    for b in range(bands):
        if scheme == "BIP":
            img_off, pix_off, line_off = b * item, bands * item, bands * item * width
        elif scheme == "BIL":
            img_off, pix_off, line_off = b * item * width, item, bands * item * width
        else:
            img_off, pix_off, line_off = b * item * width * length, item, item * width
        band = ET.SubElement(root, "VRTRasterBand", dataType=data_type, band=str(b + 1),
                             subClass="VRTRawRasterBand")
        ET.SubElement(band, "SourceFilename", relativeToVRT="1").text = os.path.basename(name)
        ET.SubElement(band, "ByteOrder").text = "LSB"
        ET.SubElement(band, "ImageOffset").text = str(img_off)
        ET.SubElement(band, "PixelOffset").text = str(pix_off)
        ET.SubElement(band, "LineOffset").text = str(line_off)
    ET.ElementTree(root).write(name + ".vrt")
