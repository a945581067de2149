"""Drop-in ``PyCuAmpcor`` class (the Python surface of contrib/PyCuAmpcor).

Same attribute names, defaults, methods and call order as the reference's
pybind11 class (contrib/PyCuAmpcor/src/PyCuAmpcor.cpp:14-101, defaults
cuAmpcorParameter.cpp:18-64):

    obj = PyCuAmpcor()
    obj.referenceImageName = ...; obj.windowSizeHeight = 64; ...
    obj.setupParams()
    obj.setConstantGrossOffset(goDown, goAcross)   # or setVaryingGrossOffset(vD, vA)
    obj.checkPixelInImageRange()                   # optional
    obj.runAmpcor()

Differences, all on error paths: where the reference calls ``exit(1)`` or
``throw;`` (cuAmpcorParameter.cpp:100-103,283-327; GDALImage.cpp:30-44) this
class raises ``ValueError`` / ``IndexError`` / ``IOError`` and the interpreter
survives. Legacy attribute names seen in callers are accepted:
``numberWindows`` (read-only), ``grossOffsetDownStatic`` and
``grossOffsetAcrossStatic``. Other unknown attributes raise
``AttributeError`` as on the pybind11 class.

Extension attributes (not in the reference): ``engineMode`` ("auto",
"staged" or "fused") and, after ``runAmpcor``, ``lastRunSeconds``.
"""
from __future__ import annotations

import ctypes as C
import operator
import time

import numpy as np
import torch

from . import _lib, raster
from .engine import MODE_AUTO, MODE_FUSED, MODE_STAGED, Engine, WindowParams

__all__ = ["PyCuAmpcor"]

# (python name, default) — int properties (PyCuAmpcor.cpp:31-76)
_INT_PROPS = {
    "algorithm": 0, "deviceID": 0, "nStreams": 1, "derampMethod": 1,
    "referenceImageHeight": 1000, "referenceImageWidth": 1000,
    "secondaryImageHeight": 1000, "secondaryImageWidth": 1000,
    "numberWindowDown": 1, "numberWindowAcross": 1,
    "windowSizeHeight": 64, "windowSizeWidth": 64,
    "mergeGrossOffset": 0, "rawDataOversamplingFactor": 2, "corrStatWindowSize": 21,
    "numberWindowDownInChunk": 1, "numberWindowAcrossInChunk": 1, "useMmap": 1,
    "halfSearchRangeAcross": 20, "halfSearchRangeDown": 20,
    "referenceStartPixelAcrossStatic": 0, "referenceStartPixelDownStatic": 0,
    "corrSurfaceOverSamplingMethod": 0, "corrSurfaceOverSamplingFactor": 16,
    "mmapSize": 1, "skipSampleDown": 64, "skipSampleAcross": 64, "corrSurfaceZoomInWindow": 16,
}
_STR_PROPS = {
    "referenceImageName": "reference.slc", "secondaryImageName": "secondary.slc",
    "offsetImageName": "DenseOffset.off", "grossOffsetImageName": "GrossOffset.off",
    "snrImageName": "snr.snr", "covImageName": "cov.cov",
}
_ALIASES = {"grossOffsetDownStatic": 0, "grossOffsetAcrossStatic": 0}
_MODES = {"auto": MODE_AUTO, "staged": MODE_STAGED, "fused": MODE_FUSED}


def _cptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PyCuAmpcor:
    """Dense amplitude cross-correlation (offset tracking) on an MI355X GPU."""

    __slots__ = ("_v", "_state")

    def __init__(self):
        object.__setattr__(self, "_v", {**_INT_PROPS, **_STR_PROPS, **_ALIASES, "engineMode": "auto"})
        object.__setattr__(self, "_state", {})

    # -- attribute surface ---------------------------------------------------
    def __getattr__(self, name):
        v = object.__getattribute__(self, "_v")
        if name in v:
            return v[name]
        if name == "numberWindows":
            return v["numberWindowDown"] * v["numberWindowAcross"]
        if name == "lastRunSeconds":
            return object.__getattribute__(self, "_state").get("seconds")
        raise AttributeError(f"'PyCuAmpcor' object has no attribute '{name}'")

    def __setattr__(self, name, value):
        v = self._v
        if name in _INT_PROPS or name in _ALIASES:
            if isinstance(value, float):
                raise TypeError(f"{name}: expected int, got float")
            v[name] = int(operator.index(value))
        elif name in _STR_PROPS:
            if isinstance(value, bytes):
                value = value.decode()
            if not isinstance(value, str):
                raise TypeError(f"{name}: expected str")
            v[name] = value
        elif name == "engineMode":
            if value not in _MODES:
                raise ValueError(f"engineMode must be one of {sorted(_MODES)}")
            v[name] = value
        elif name == "numberWindows":
            raise AttributeError("numberWindows is read-only (numberWindowDown*numberWindowAcross)")
        else:
            raise AttributeError(f"'PyCuAmpcor' object has no attribute '{name}'")

    # -- parameters ------------------------------------------------------------
    def _window_params(self) -> WindowParams:
        v = self._v
        return WindowParams(
            window_h=v["windowSizeHeight"], window_w=v["windowSizeWidth"],
            hs_down=v["halfSearchRangeDown"], hs_across=v["halfSearchRangeAcross"],
            raw_ovs=v["rawDataOversamplingFactor"], zoom=v["corrSurfaceZoomInWindow"],
            surf_ovs=v["corrSurfaceOverSamplingFactor"], surf_method=v["corrSurfaceOverSamplingMethod"],
            corr_stat=v["corrStatWindowSize"], deramp=v["derampMethod"], algorithm=v["algorithm"],
            image_complex=1)

    def setupParams(self):
        """cuAmpcorParameter::setupParameters (cuAmpcorParameter.cpp:70-109)."""
        v = self._v
        nd, na = v["numberWindowDown"], v["numberWindowAcross"]
        if nd * na <= 0 or nd <= 0 or na <= 0:
            raise ValueError(f"Incorrect number of windows! ({nd}, {na})")
        if v["numberWindowDownInChunk"] <= 0 or v["numberWindowAcrossInChunk"] <= 0:
            raise ValueError("numberWindow*InChunk must be positive")
        p = self._window_params()
        # the zoom clamp of setupParameters is visible through the property (cuAmpcorParameter.cpp:80-83)
        zoom = min(p.zoom, min(p.hs_across, p.hs_down) * 2 * p.raw_ovs)
        v["corrSurfaceZoomInWindow"] = zoom
        hz = zoom // (2 * p.raw_ovs)
        st = self._state
        st.clear()
        st["sizes"] = dict(
            zoom=zoom, half_zoom=hz,
            search_h=p.window_h + 2 * p.hs_down, search_w=p.window_w + 2 * p.hs_across)
        st["chunks"] = (-(-nd // v["numberWindowDownInChunk"]), -(-na // v["numberWindowAcrossInChunk"]))

    def _need_setup(self):
        if "sizes" not in self._state:
            raise RuntimeError("setupParams() must be called first")

    def _set_starts(self, g_down, g_across):
        """cuAmpcorParameter::setStartPixels (cuAmpcorParameter.cpp:167-205),
        through the C ABI's ``ampcor_start_pixels``."""
        v = self._v
        nd, na = v["numberWindowDown"], v["numberWindowAcross"]
        n = nd * na
        lib = _lib.load()
        p = self._window_params().to_c()
        out = [np.empty(n, np.int32) for _ in range(4)]
        varying = np.ndim(g_down) > 0
        if varying:
            gd = np.ascontiguousarray(np.broadcast_to(np.asarray(g_down), (n,)), np.int32)
            ga = np.ascontiguousarray(np.broadcast_to(np.asarray(g_across), (n,)), np.int32)
            g0 = (0, 0)
        else:
            gd = ga = None
            g0 = (int(g_down), int(g_across))
        code = lib.ampcor_start_pixels(C.byref(p), nd, na, v["referenceStartPixelDownStatic"],
                                       v["referenceStartPixelAcrossStatic"], v["skipSampleDown"],
                                       v["skipSampleAcross"], g0[0], g0[1], _cptr(gd), _cptr(ga), *map(_cptr, out))
        _lib.check(code, None)
        st = self._state
        st["ref_d"], st["ref_a"], st["sec_d"], st["sec_a"] = (a.astype(np.int64) for a in out)
        st["gross_d"] = gd.astype(np.int64) if varying else np.full(n, g0[0], np.int64)
        st["gross_a"] = ga.astype(np.int64) if varying else np.full(n, g0[1], np.int64)

    def setConstantGrossOffset(self, goDown, goAcross):
        self._need_setup()
        self._set_starts(int(operator.index(goDown)), int(operator.index(goAcross)))

    def setVaryingGrossOffset(self, vD, vA):
        self._need_setup()
        vD = np.asarray(vD).astype(np.int64).ravel()
        vA = np.asarray(vA).astype(np.int64).ravel()
        n = self.numberWindows
        if vD.size < n or vA.size < n:
            raise ValueError(f"varying gross offsets need {n} values per direction, got {vD.size}/{vA.size}")
        self._set_starts(vD[:n], vA[:n])

    def checkPixelInImageRange(self):
        """cuAmpcorParameter::checkPixelInImageRange (cuAmpcorParameter.cpp:275-331),
        through the C ABI's ``ampcor_check_range``: strict end test
        (``start + size >= dim`` fails). Raises ``IndexError`` naming the first
        failing window (row, col), the test and the offending pixel."""
        st, v = self._state, self._v
        if "ref_d" not in st:
            raise RuntimeError("set a gross offset first")
        sz = st["sizes"]
        arrs = [np.ascontiguousarray(st[k], np.int32) for k in ("ref_d", "ref_a", "sec_d", "sec_a")]
        p = self._window_params().to_c()
        bad, which = C.c_int32(), C.c_int32()
        lib = _lib.load()
        code = lib.ampcor_check_range(C.byref(p), *map(_cptr, arrs), len(arrs[0]), v["referenceImageHeight"],
                                      v["referenceImageWidth"], v["secondaryImageHeight"], v["secondaryImageWidth"],
                                      C.byref(bad), C.byref(which))
        if code == _lib.AMPCOR_OK:
            return
        if code != -2:
            _lib.check(code, None)
        i, k = bad.value, which.value
        start = int(arrs[k // 4 * 2 + (k % 2)][i])
        size = ((v["windowSizeHeight"], v["windowSizeWidth"]), (sz["search_h"], sz["search_w"]))[k // 4][k % 2]
        pixel = start if (k % 4) < 2 else start + size
        kind = ("Reference", "Secondary")[k // 4]
        edge = ("start", "end")[(k % 4) // 2]
        axis = ("Down", "Across")[k % 2]
        row, col = divmod(i, v["numberWindowAcross"])
        raise IndexError(f"{kind} Window {edge} pixel out of range in {axis}, window ({row},{col}), pixel {pixel}")

    # -- run -----------------------------------------------------------------------
    def _engine(self, dev_index):
        st = self._state
        eng = st.get("engine")
        if eng is None or eng.device != dev_index:
            eng = Engine(dev_index)
            st["engine"] = eng
        return eng

    def runAmpcor(self):
        """cuAmpcorController::runAmpcor (cuAmpcorController.cpp:37-178): open the two
        rasters, run every window, write offset / gross offset / snr / cov files."""
        self._need_setup()
        st, v = self._state, self._v
        if "ref_d" not in st:
            raise RuntimeError("setConstantGrossOffset or setVaryingGrossOffset must be called first")
        if not torch.cuda.is_available():
            raise RuntimeError("runAmpcor needs a HIP device (the engine has no CPU path)")
        ref_np = raster.open_raster(v["referenceImageName"], v["referenceImageHeight"], v["referenceImageWidth"])
        sec_np = raster.open_raster(v["secondaryImageName"], v["secondaryImageHeight"], v["secondaryImageWidth"])
        cplx = np.iscomplexobj(ref_np)
        if cplx != np.iscomplexobj(sec_np):
            raise raster.RasterError("reference and secondary rasters must both be complex or both real")
        t0 = time.perf_counter()
        # row bands of the rasters are staged and uploaded while earlier bands of
        # windows run (the reference loads and runs chunk by chunk in turn)
        eng, mode, chunk = self._prepared_engine(v["deviceID"], cplx)
        out = eng.run_windows_from_host(ref_np, sec_np, st["ref_d"], st["ref_a"], st["sec_d"], st["sec_a"],
                                        mode=mode, chunk=chunk)
        st["seconds"] = time.perf_counter() - t0
        nd, na = v["numberWindowDown"], v["numberWindowAcross"]
        offset = out["offset"].reshape(nd, na, 2)
        gross = np.stack([st["gross_d"], st["gross_a"]], axis=1).astype(np.float32).reshape(nd, na, 2)
        if v["mergeGrossOffset"]:
            offset = offset + gross
        raster.write_raw(v["offsetImageName"], offset)
        raster.write_raw(v["grossOffsetImageName"], gross)
        raster.write_raw(v["snrImageName"], out["snr"].reshape(nd, na))
        raster.write_raw(v["covImageName"], out["cov"].reshape(nd, na, 3))
        st["result"] = dict(offset=offset, gross=gross, snr=out["snr"].reshape(nd, na),
                            cov=out["cov"].reshape(nd, na, 3))
        return None

    def _prepared_engine(self, device, cplx):
        """The engine for ``device`` set up for this object's parameters, with the
        mode and chunk that ``engineMode`` and the chunk attributes select."""
        v = self._v
        eng = self._engine(device)
        p = self._window_params()
        p.image_complex = 1 if cplx else 0
        eng.setup(p)
        mode = _MODES[v["engineMode"]]
        chunk = v["numberWindowDownInChunk"] * v["numberWindowAcrossInChunk"] if mode == MODE_STAGED else 0
        return eng, mode, chunk if chunk > 1 else 0

    def run_arrays(self, ref, sec, debug=False):
        """Run on device-resident images (torch tensors); returns device tensors.

        The windows are those set by ``set*GrossOffset``; the engine checks
        that every chip lies inside its image."""
        self._need_setup()
        st = self._state
        dev = ref.device
        eng, mode, chunk = self._prepared_engine(dev.index or 0, ref.is_complex())
        starts = [torch.as_tensor(np.ascontiguousarray(st[k], dtype=np.int32)).to(dev)
                  for k in ("ref_d", "ref_a", "sec_d", "sec_a")]
        return eng.run_windows(ref, sec, *starts, mode=mode, chunk=chunk, debug=debug)
