"""Bit-exact 2-bit unpack on the GPU (kernel.c:388-426, layout kernel.c:195-200).

Two routes through the C ABI, both against the oracle's unpack (itself pinned
by brute-force bit extraction in test_oracle.py) with np.array_equal:

* the kernel.h entry points `unpack` and `unpackReal2bit` (kernel.h:67-70) at
  even and odd sample offsets, the buffer start, the window edges and the
  last offsets a subint allows (maxoffset, kernel.c:730), for nchan 256 ...
  16384;
* the levels the fused F-stage decodes inside its pass 0 (LDS byte window
  recoded to code pairs + level table), read through the F-stage's level tap,
  for every radix plan and every output precision, with delay models whose
  offsets are odd and even and hit both edge clamps (kernel.c:742-756).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, engine, synth

pytestmark = pytest.mark.gpu

LO, BW = 1.65e9, 32e6


def _kernelh():
    lib = engine.load_library()
    P, I = C.c_void_p, C.c_int
    lib.unpack.argtypes = [P, P, I, I, I]
    lib.unpack.restype = None
    lib.unpackReal2bit.argtypes = [P, P, I, I]
    lib.unpackReal2bit.restype = None
    return lib


def _call(fn, data, offset, n, *extra):
    u0 = np.full(n, np.nan, np.complex64)
    u1 = np.full(n, np.nan, np.complex64)
    ptrs = (C.c_void_p * 2)(u0.ctypes.data, u1.ctypes.data)
    fn(C.c_void_p(data.ctypes.data), ptrs, offset, *extra)
    return u0, u1


@pytest.mark.parametrize("nchan", [256, 1024, 4096, 16384])
def test_kernelh_unpack_bit_exact(nchan):
    lib = _kernelh()
    F = 2 * nchan
    numffts = 5
    maxoffset = (numffts - 1) * F
    data = synth.random_bytes(1, (numffts + 1) * F // 2, 900 + nchan)[0]
    offsets = sorted({0, 1, 2, 3, 5, F // 2 - 1, F - 1, F, F + 1, 2 * F + 7, maxoffset - 2, maxoffset - 1,
                      maxoffset, maxoffset + 1})
    for o in offsets:
        u0, u1 = _call(lib.unpack, data, o, F, 2, F)
        l0, l1 = orc.unpack(data, o, F)
        assert np.array_equal(u0.real, l0) and np.array_equal(u1.real, l1), (nchan, o)
        assert not np.any(u0.imag) and not np.any(u1.imag), (nchan, o)
    # unpackReal2bit with sample counts that are not an FFT length
    for o, n in ((0, 1), (1, 1), (1, 37), (6, 64), (F - 3, 101)):
        u0, u1 = _call(lib.unpackReal2bit, data, o, n, n)
        l0, l1 = orc.unpack(data, o, n)
        assert np.array_equal(u0.real, l0) and np.array_equal(u1.real, l1), (nchan, o, n)
        assert not np.any(u0.imag) and not np.any(u1.imag)


def _edge_model(nant, Ts, nf, F, seed):
    """Delay models with net delays of a few samples of either sign, varying
    fast enough that some FFT offsets are odd, some even, and the first and
    last FFTs of some antennas land on the edge clamps o = -1 and
    o = maxoffset + 1 (or fall outside)."""
    rng = np.random.default_rng(seed)
    c2 = rng.uniform(-2e-6, 2e-6, nant)
    spread = np.linspace(-2.6, 2.6, nant)                  # net delay at FFT 0, samples
    rate = rng.uniform(-1.5, 1.5, nant) * Ts               # samples of drift per FFT
    delays = np.stack([np.zeros(nant), rate, c2], axis=1)
    fileoff = c2 - spread * Ts
    return delays, fileoff


@pytest.mark.parametrize("prec", ["f32", "f16", "f8"])
@pytest.mark.parametrize("nchan", [256, 512, 1024, 2048, 4096, 8192, 16384])
def test_fstage_decoded_levels_bit_exact(nchan, prec):
    nant = 5
    F = 2 * nchan
    p = orc.plan(nchan, LO, BW)
    with FxEngine(nant, nchan, LO, BW, prec=prec, max_ffts=64, max_bytes_per_ant=65 * F // 2) as e:
        nf = e.granule                                      # 8, 16 or 64 FFTs
        delays, fileoff = _edge_model(nant, p.Ts, nf, F, nchan)
        data = synth.random_bytes(nant, (nf + 1) * F // 2, 77 + nchan)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.enable_tap(True, levels=True)
        it = np.arange(nf, dtype=np.int32)
        seen_odd = seen_even = 0
        for plan in range(e.nplans()):
            e.set_plan(plan)
            e.fstage(it)
            e.sync()
            lv = e.read_tap_levels(nf)
            tab = e.read_delay_table(nf)
            for k in range(nf):
                for j in range(nant):
                    m, a, b = orc.station_delay(delays[j], k)
                    v, o, fd = orc.offset(m, fileoff[j], p.Ts, k, F, (nf - 1) * F)
                    assert (tab[k, j]["valid"], tab[k, j]["offset"]) == (v, o)
                    if not v:
                        continue
                    seen_odd += o & 1
                    seen_even += 1 - (o & 1)
                    l0, l1 = orc.unpack(data[j], o, F)
                    assert np.array_equal(lv[k, j, 0], l0), (plan, k, j, o)
                    assert np.array_equal(lv[k, j, 1], l1), (plan, k, j, o)
        assert seen_odd and seen_even
