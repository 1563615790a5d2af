"""ctypes bindings of the native FX engine (include/difxfx.h).

The engine is libdifxfx.so built in-tree by difx_tests_amd.build; nothing
here computes visibilities on the CPU.  Loading fails loudly when the library
is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIBPATH = os.path.join(LIBDIR, "libdifxfx.so")

PREC_F32, PREC_F16, PREC_F8, PREC_F32V = 0, 1, 2, 3
PREC_NAMES = {"f32": PREC_F32, "f16": PREC_F16, "f8": PREC_F8, "f32v": PREC_F32V}

# Every symbol include/difxfx.h declares.
NATIVE_SYMBOLS = [
    "fx_create", "fx_destroy", "fx_last_error", "fx_fft_granule", "fx_set_model", "fx_upload_data",
    "fx_upload_data_strided", "fx_device_data", "fx_fstage", "fx_xstage", "fx_process", "fx_reset",
    "fx_read_vis", "fx_read_delay_table", "fx_enable_tap", "fx_read_tap", "fx_set_fp8_scales", "fx_sync",
    "fx_set_timing", "fx_reset_timing", "fx_get_timing", "fx_get_geometry", "fx_set_option", "fx_get_option",
    "fx_batch_create", "fx_batch_destroy", "fx_batch_process", "fx_batch_sync", "fx_batch_captures",
    "fx_autotune", "fx_fstage_ex", "fx_ingest_host", "fx_ingest_submit", "fx_ingest_wait", "fx_read_vis_ex",
    "fx_save_vis", "fx_batch_process_ex", "fx_read_tap_levels", "fx_subint", "fx_get_graph_timing",
    "fx_process_noalloc", "fx_update_fp8_scales",
]
VIS_NORMALIZE = 1
TAP_SPECTRA, TAP_LEVELS = 1, 2
VIS_TEXT, VIS_BINARY = 0, 1
OPT_XTILE, OPT_FPLAN, OPT_NPLANS, OPT_XBATCH, OPT_GRAPH, OPT_FP8_AUTOSCALE = 1, 2, 3, 4, 5, 6
BATCH_LAUNCH, BATCH_GRAPH = 0, 1
# Every symbol include/kernel.h declares (kernel.h:61-134 of the reference).
KERNELH_SYMBOLS = [
    "initLUT2bitReal", "init_FxKernel", "unpack", "getStationDelay", "unpackReal2bit", "fringerotate", "dofft",
    "fracSampleCorrect", "conjChannels", "saveVisibilities", "startTiming", "saveLog", "endTiming",
    "processAntennas", "processNormalize", "processBaseline", "processAntennasAndBaseline",
    "processAntennasAndBaselinePara", "resetBeforeProcess", "manageAntennasAndBaseline", "processBaselinePara",
    "merge", "unpackImpl", "fringeRotateImpl", "doFFTImpl", "fracSampleCorrectImpl", "conjChannelsImpl",
    "stationDelayAndOffset",
]

DELAY_DTYPE = np.dtype([("a", "<f8"), ("b", "<f8"), ("fd", "<f8"), ("offset", "<i4"), ("valid", "<i4")])


class FxError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [("nant", C.c_int), ("nchan", C.c_int), ("lo", C.c_double), ("bw", C.c_double),
                ("prec", C.c_int), ("max_ffts", C.c_int), ("max_bytes_per_ant", C.c_longlong),
                ("iscomplex", C.c_int)]


_lib = None


def load_library(path: str = LIBPATH) -> C.CDLL:
    """Load libdifxfx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FxError(f"{path} is missing: run `python -m difx_tests_amd.build` first")
    lib = C.CDLL(path)
    P = C.c_void_p
    I = C.c_int
    lib.fx_create.argtypes = [C.POINTER(_Params), C.POINTER(P)]
    lib.fx_destroy.argtypes = [P]
    lib.fx_destroy.restype = None
    lib.fx_last_error.restype = C.c_char_p
    lib.fx_fft_granule.argtypes = [P]
    lib.fx_set_model.argtypes = [P, P, P]
    lib.fx_upload_data.argtypes = [P, P, C.c_longlong]
    lib.fx_upload_data_strided.argtypes = [P, P, C.c_longlong, C.c_longlong]
    lib.fx_device_data.argtypes = [P, C.POINTER(P), C.POINTER(C.c_longlong)]
    lib.fx_fstage.argtypes = [P, I, P, I, I, C.POINTER(I)]
    lib.fx_xstage.argtypes = [P, I, I]
    lib.fx_process.argtypes = [P, I, P, I]
    lib.fx_reset.argtypes = [P]
    lib.fx_read_vis.argtypes = [P, P, P, P, P]
    lib.fx_read_delay_table.argtypes = [P, I, P]
    lib.fx_enable_tap.argtypes = [P, I]
    lib.fx_read_tap.argtypes = [P, I, P]
    lib.fx_read_tap_levels.argtypes = [P, I, P]
    lib.fx_subint.argtypes = [P, I, I, P, I, I, I, I, C.POINTER(I)]
    lib.fx_get_graph_timing.argtypes = [P, P, P]
    lib.fx_set_fp8_scales.argtypes = [P, P]
    lib.fx_update_fp8_scales.argtypes = [P]
    lib.fx_sync.argtypes = [P]
    lib.fx_set_timing.argtypes = [P, I]
    lib.fx_reset_timing.argtypes = [P]
    lib.fx_get_timing.argtypes = [P, P, P]
    lib.fx_get_geometry.argtypes = [P, P]
    lib.fx_set_option.argtypes = [P, I, I]
    lib.fx_get_option.argtypes = [P, I, C.POINTER(I)]
    lib.fx_batch_create.argtypes = [P, I, C.POINTER(P)]
    lib.fx_batch_destroy.argtypes = [P]
    lib.fx_batch_destroy.restype = None
    lib.fx_batch_process.argtypes = [P, P, P, I, I]
    lib.fx_batch_process_ex.argtypes = [P, P, P, P, I, I]
    lib.fx_batch_sync.argtypes = [P]
    lib.fx_batch_captures.argtypes = [P]
    lib.fx_autotune.argtypes = [P, I, I, C.c_char_p, P]
    lib.fx_fstage_ex.argtypes = [P, I, I, P, I, I, I, C.POINTER(I)]
    lib.fx_ingest_host.argtypes = [P, I, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_longlong)]
    lib.fx_ingest_submit.argtypes = [P, I, C.c_longlong]
    lib.fx_ingest_wait.argtypes = [P, I]
    lib.fx_read_vis_ex.argtypes = [P, I, P, P, P, P]
    lib.fx_save_vis.argtypes = [P, C.c_char_p, I, I]
    _lib = lib
    return lib


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class FxEngine:
    """One subband on the GPU: delay model, packed input, F-stage, X-engine."""

    def __init__(self, nant: int, nchan: int, lo: float, bw: float, prec: int | str = PREC_F32,
                 max_ffts: int = 64, max_bytes_per_ant: int = 0, complex8: bool = False):
        """complex8: complex 8-bit input (4 bytes per time sample: pol 0 re, im,
        pol 1 re, im as int8), F = nchan; otherwise real 2-bit, F = 2*nchan."""
        self.lib = load_library()
        if isinstance(prec, str):
            prec = PREC_NAMES[prec]
        self.nant, self.nchan, self.lo, self.bw, self.prec = nant, nchan, float(lo), float(bw), prec
        self.complex8 = bool(complex8)
        prm = _Params(nant, nchan, float(lo), float(bw), prec, max_ffts, max_bytes_per_ant, 1 if complex8 else 0)
        h = C.c_void_p()
        self._check(self.lib.fx_create(C.byref(prm), C.byref(h)))
        self.h = h
        geo = np.zeros(11, np.int32)
        self._check(self.lib.fx_get_geometry(self.h, _ptr(geo)))
        (self.F, self.stride, self.sub, self.nstep, self.nfstep, self.S, self.Spad, self.nblk, self.nbp,
         self.G, self.granule) = [int(v) for v in geo]
        self.nbl = nant * (nant - 1) // 2

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise FxError(f"libdifxfx error {rc}: {self.lib.fx_last_error().decode()}")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.fx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- inputs
    def set_model(self, delays: np.ndarray, fileoffsets: np.ndarray) -> None:
        d = np.ascontiguousarray(delays, dtype=np.float64).reshape(self.nant, 3)
        o = np.ascontiguousarray(fileoffsets, dtype=np.float64).reshape(self.nant)
        self._check(self.lib.fx_set_model(self.h, _ptr(d), _ptr(o)))

    def upload(self, data: np.ndarray) -> None:
        """data: uint8 [nant][nbytes] packed 2-bit samples."""
        d = np.ascontiguousarray(data, dtype=np.uint8)
        assert d.ndim == 2 and d.shape[0] == self.nant
        self._check(self.lib.fx_upload_data_strided(self.h, _ptr(d), d.shape[1], d.shape[1]))

    # ---- pipeline
    def fstage(self, iteration, split: int = 1, fft0: int = 0) -> int:
        it = np.ascontiguousarray(iteration, dtype=np.int32)
        ns = C.c_int()
        self._check(self.lib.fx_fstage(self.h, len(it), _ptr(it), split, fft0, C.byref(ns)))
        return ns.value

    def fstage_ex(self, iteration, slot: int = 0, split: int = 1, origin: int = 0, fft0: int = 0) -> int:
        """F-stage on input slot `slot` of a buffer whose first FFT is `origin`
        (global FFT indices in `iteration`)."""
        it = np.ascontiguousarray(iteration, dtype=np.int32)
        ns = C.c_int()
        self._check(self.lib.fx_fstage_ex(self.h, slot, len(it), _ptr(it), split, origin, fft0, C.byref(ns)))
        return ns.value

    # ---- streaming ingest
    def ingest_host(self, slot: int) -> np.ndarray:
        """Pinned host buffer of an input slot as uint8 [nant, stride]."""
        ptr = C.POINTER(C.c_uint8)()
        stride = C.c_longlong()
        self._check(self.lib.fx_ingest_host(self.h, slot, C.byref(ptr), C.byref(stride)))
        return np.ctypeslib.as_array(ptr, shape=(self.nant, stride.value))

    def ingest_submit(self, slot: int, nbytes: int) -> None:
        self._check(self.lib.fx_ingest_submit(self.h, slot, nbytes))

    def ingest_wait(self, slot: int) -> None:
        self._check(self.lib.fx_ingest_wait(self.h, slot))

    def subint(self, iteration, slot: int = 0, split: int = 1, origin: int = 0, fft0: int = 0,
               xslots: int = 0) -> int:
        """One subint as one graph launch: delay table, F-stage at fft0 and,
        when xslots > 0, the X-engine over slots [0, xslots)."""
        it = np.ascontiguousarray(iteration, dtype=np.int32)
        ns = C.c_int()
        self._check(self.lib.fx_subint(self.h, slot, len(it), _ptr(it), split, origin, fft0, xslots, C.byref(ns)))
        return ns.value

    def set_graph(self, on: bool) -> None:
        """fx_subint / fx_process as recorded graphs (default) or plain launches."""
        self._check(self.lib.fx_set_option(self.h, OPT_GRAPH, 1 if on else 0))

    def graph_timing(self):
        """(milliseconds, launches) of fx_subint graph launches since reset_timing."""
        ms = C.c_double()
        n = C.c_int()
        self._check(self.lib.fx_get_graph_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def xstage(self, fft0: int, nslots: int) -> None:
        self._check(self.lib.fx_xstage(self.h, fft0, nslots))

    def process(self, iteration, split: int = 1) -> None:
        it = np.ascontiguousarray(iteration, dtype=np.int32)
        self._check(self.lib.fx_process(self.h, len(it), _ptr(it), split))

    def reset(self) -> None:
        self._check(self.lib.fx_reset(self.h))

    def sync(self) -> None:
        self._check(self.lib.fx_sync(self.h))

    # ---- outputs
    def read_vis(self, normalize: bool = False):
        """(cross [nbl,4,nchan] c64, counts [nbl], autos [nant,4,nchan] c64, autocounts [nant]);
        normalize divides by the valid-FFT counts on the device."""
        cross = np.zeros((max(self.nbl, 1), 4, self.nchan), np.complex64)
        autos = np.zeros((self.nant, 4, self.nchan), np.complex64)
        cc = np.zeros(max(self.nbl, 1), np.int32)
        ac = np.zeros(self.nant, np.int32)
        self._check(self.lib.fx_read_vis_ex(self.h, VIS_NORMALIZE if normalize else 0, _ptr(cross), _ptr(cc),
                                            _ptr(autos), _ptr(ac)))
        return cross[: self.nbl], cc[: self.nbl], autos, ac

    def save_vis(self, path: str, fmt: str = "binary", normalize: bool = True) -> None:
        """Writes the visibilities: fmt "binary" (DIFXFXV1) or "text" (vis.out)."""
        self._check(self.lib.fx_save_vis(self.h, path.encode(), VIS_BINARY if fmt == "binary" else VIS_TEXT,
                                         VIS_NORMALIZE if normalize else 0))

    def read_delay_table(self, nffts: int) -> np.ndarray:
        out = np.zeros(nffts * self.nant, DELAY_DTYPE)
        self._check(self.lib.fx_read_delay_table(self.h, nffts, _ptr(out)))
        return out.reshape(nffts, self.nant)

    def enable_tap(self, on: bool = True, levels: bool = False) -> None:
        """FP32 spectra of every F-stage FFT; with levels also the 2-bit levels
        the fused F-stage decoded (real input)."""
        mask = (TAP_SPECTRA | (TAP_LEVELS if levels else 0)) if on else 0
        self._check(self.lib.fx_enable_tap(self.h, mask))

    def read_tap_levels(self, nffts: int) -> np.ndarray:
        """float32 [nffts, nant, 2, F] decoded levels of the valid FFTs."""
        out = np.full((nffts, self.nant, 2, self.F), np.nan, np.float32)
        self._check(self.lib.fx_read_tap_levels(self.h, nffts, _ptr(out)))
        return out

    def read_tap(self, nffts: int) -> np.ndarray:
        out = np.zeros((nffts, self.nant, 2, self.nchan), np.complex64)
        self._check(self.lib.fx_read_tap(self.h, nffts, _ptr(out)))
        return out

    def set_fp8_scales(self, scales: np.ndarray) -> None:
        s = np.ascontiguousarray(scales, dtype=np.float32).reshape(self.nchan, 2 * self.nant)
        self._check(self.lib.fx_set_fp8_scales(self.h, _ptr(s)))

    def update_fp8_scales(self) -> None:
        """FP8 scales from the autocorrelations accumulated since reset()."""
        self._check(self.lib.fx_update_fp8_scales(self.h))

    def set_fp8_autoscale(self, on: bool = True) -> None:
        """FP8 scales follow the accumulated autocorrelations after every X-engine launch."""
        self._check(self.lib.fx_set_option(self.h, OPT_FP8_AUTOSCALE, 1 if on else 0))

    # ---- tuning
    def _get(self, opt: int) -> int:
        v = C.c_int()
        self._check(self.lib.fx_get_option(self.h, opt, C.byref(v)))
        return v.value

    def set_xtile(self, sb: int) -> None:
        """X-engine workgroup tile (1, 2 or 4 blocks per panel side; 0: automatic)."""
        self._check(self.lib.fx_set_option(self.h, OPT_XTILE, sb))

    def xtile(self) -> int:
        return self._get(OPT_XTILE)

    def set_plan(self, v: int) -> None:
        """F-stage radix plan, 0 .. nplans() - 1."""
        self._check(self.lib.fx_set_option(self.h, OPT_FPLAN, v))

    def plan(self) -> int:
        return self._get(OPT_FPLAN)

    def nplans(self) -> int:
        return self._get(OPT_NPLANS)

    def xbatch(self) -> int:
        """Subints per X-engine launch recommended by the autotuner."""
        return self._get(OPT_XBATCH)

    def autotune(self, numffts: int, reps: int = 3, cache: str | None = None) -> dict:
        """Times the radix plans, X-engine tiles and subints per X launch on the
        uploaded input; keeps the fastest and zeroes the visibilities."""
        rep = np.zeros(5, np.float64)
        self._check(self.lib.fx_autotune(self.h, numffts, reps, cache.encode() if cache else None, _ptr(rep)))
        return {"plan": int(rep[0]), "xtile": int(rep[1]), "xbatch": int(rep[2]), "fstage_ms": float(rep[3]),
                "xstage_ms_per_subint": float(rep[4])}

    # ---- timing
    def set_timing(self, on: bool = True) -> None:
        self._check(self.lib.fx_set_timing(self.h, 1 if on else 0))

    def reset_timing(self) -> None:
        self._check(self.lib.fx_reset_timing(self.h))

    def timing(self):
        """({delay, fstage, counts, xstage} ms summed, {fstage, xstage} launch counts)."""
        ms = np.zeros(4, np.float64)
        nl = np.zeros(2, np.int32)
        self._check(self.lib.fx_get_timing(self.h, _ptr(ms), _ptr(nl)))
        return dict(zip(("delay", "fstage", "counts", "xstage"), ms.tolist())), nl.tolist()


class FxBatch:
    """Several engines (one per subband) run one subint at a time, as one
    hipGraph launch (mode "graph") or as plain launches one subband after
    another (mode "launch")."""

    def __init__(self, engines):
        self.lib = load_library()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        self._check(self.lib.fx_batch_create(arr, len(self.engines), C.byref(h)))
        self.h = h

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise FxError(f"libdifxfx error {rc}: {self.lib.fx_last_error().decode()}")

    def process(self, iterations, split: int = 1, mode: str = "graph", origins=None) -> None:
        its = [np.ascontiguousarray(i, dtype=np.int32) for i in iterations]
        assert len(its) == len(self.engines)
        n = np.array([len(i) for i in its], np.int32)
        ptrs = (C.c_void_p * len(its))(*[i.ctypes.data for i in its])
        m = BATCH_GRAPH if mode == "graph" else BATCH_LAUNCH
        org = None if origins is None else np.ascontiguousarray(origins, dtype=np.int32)
        self._check(self.lib.fx_batch_process_ex(self.h, _ptr(n), ptrs, None if org is None else _ptr(org), split,
                                                 m))

    def sync(self) -> None:
        self._check(self.lib.fx_batch_sync(self.h))

    @property
    def captures(self) -> int:
        return int(self.lib.fx_batch_captures(self.h))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.fx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
