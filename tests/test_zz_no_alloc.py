"""fx_process_noalloc (the allocation-free entry point) against fx_process.

Two calls in a row through the C ABI on the same GPU, the second with a
smaller input: a 16-antenna x 4096-channel subint of 100 FFTs (more than the
scratch array holds at once, so it runs in FFT chunks) and then a 4-antenna
x 512-channel subint of 13 FFTs (one chunk).  Inputs and outputs are device
buffers the test allocates (hipMalloc of the engine's own HIP runtime); the
entry point allocates nothing.  Each call's
visibilities and counts are compared with a context's fx_process +
fx_read_vis on the same inputs: counts exact, every (baseline, product)
within 1e-5 (FP32, FP16 and FP8 X-engines; the chunked call adds the same
products in a different grouping, the one-chunk call is bit-identical).  A
workload beyond the array is refused before any launch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, engine, synth
from difx_tests_amd.engine import PREC_NAMES, _Params

pytestmark = pytest.mark.gpu

LO, BW = 1.65e9, 32e6


def _lib():
    lib = engine.load_library()
    P, I = C.c_void_p, C.c_int
    lib.fx_process_noalloc.argtypes = [C.POINTER(_Params), P, C.c_longlong, P, P, P, I, I, P, P, P, P, P]
    return lib


def _case(nant, nchan, nf, seed):
    p = orc.plan(nchan, LO, BW)
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-3e-10, 3e-10, nant),
                       rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-1.4, 1.4, nant) * p.Ts
    data = synth.make_data(nant, (nf + 1) * 2 * nchan, seed)
    return data, delays, fileoff, np.arange(nf, dtype=np.int32)


class _Hip:
    """Device buffers through the HIP runtime the engine itself links
    (libamdhip64 of /opt/rocm): a second runtime in the process, such as the
    one a PyTorch wheel bundles, cannot open the GPU once this one has."""

    def __init__(self):
        self.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, C.c_void_p(a.ctypes.data), a.nbytes, 1) == 0      # host to device
        return p.value

    def get(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(lib, prec, data, delays, fileoff, it, nchan):
    nant, nbytes = data.shape
    stride = (nbytes + 16 + 255) // 256 * 256
    buf = np.zeros((nant, stride), np.uint8)
    buf[:, :nbytes] = data
    nbl = nant * (nant - 1) // 2
    hip = _Hip()
    try:
        d_data = hip.put(buf)
        d_delays = hip.put(np.asarray(delays, np.float64))
        d_off = hip.put(np.asarray(fileoff, np.float64))
        d_it = hip.put(np.asarray(it, np.int32))
        cross0 = np.full((nbl, 4, nchan), np.nan, np.complex64)
        autos0 = np.full((nant, 4, nchan), np.nan, np.complex64)
        cc0 = np.full(nbl, -1, np.int32)
        ac0 = np.full(nant, -1, np.int32)
        cross, autos, cc, ac = hip.put(cross0), hip.put(autos0), hip.put(cc0), hip.put(ac0)
        prm = _Params(nant, nchan, LO, BW, PREC_NAMES[prec], len(it), 0, 0)
        rc = lib.fx_process_noalloc(C.byref(prm), d_data, stride, d_delays, d_off, d_it, len(it), 1, cross, cc, autos,
                                    ac, None)
        assert rc == 0, lib.fx_last_error().decode()
        return hip.get(cross, cross0), hip.get(cc, cc0), hip.get(autos, autos0), hip.get(ac, ac0)
    finally:
        hip.free()


def _context(prec, data, delays, fileoff, it, nchan):
    nant = data.shape[0]
    with FxEngine(nant, nchan, LO, BW, prec=prec, max_ffts=len(it), max_bytes_per_ant=data.shape[1]) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        return e.read_vis()


def _close(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    den = np.linalg.norm(b, axis=-1)
    return float((num / np.maximum(den, 1e-30)).max())


@pytest.mark.parametrize("prec", ["f32", "f16", "f8"])
def test_noalloc_matches_fx_process_two_calls(prec):
    lib = _lib()
    cases = [(16, 4096, 100, 61), (4, 512, 13, 62)]          # the second call is smaller
    for nant, nchan, nf, seed in cases:
        data, delays, fileoff, it = _case(nant, nchan, nf, seed)
        got = _noalloc(lib, prec, data, delays, fileoff, it, nchan)
        want = _context(prec, data, delays, fileoff, it, nchan)
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[3], want[3]), (nant, nchan)
        assert (want[1] > 0).all()
        assert _close(got[0], want[0]) <= 1e-5, (prec, nant, nchan)
        assert _close(got[2], want[2]) <= 1e-5, (prec, nant, nchan)
        if nant == 4:                                          # one chunk: the same launches as fx_process
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[2], want[2])


def test_noalloc_refuses_oversized_workload():
    lib = _lib()
    prm = _Params(256, 4096, LO, BW, PREC_NAMES["f16"], 16, 0, 0)
    dummy = C.c_void_p(256)                                    # never dereferenced: refused on the host
    rc = lib.fx_process_noalloc(C.byref(prm), dummy, 1 << 20, dummy, dummy, dummy, 16, 1, None, None, None, None,
                                None)
    assert rc == -3 and b"scratch" in lib.fx_last_error()
