"""Several subints in one delay / F-stage launch (the benchmark's default
step): one fx_fstage call over the FFT indices of consecutive subints of a
buffer that holds them all, against one call per subint with the buffer's
`split`, and against the oracle.

The delay tables, the edge clamps and every FFT's spectra are the same either
way (each FFT is formed alone; the clamps act at the buffer's ends), so the
FP32 spectra read through the tap are compared bit for bit.  The X-engine
sees the same rows in other K chunks (no padding slots between the subints),
so its sums differ only in their order: visibilities within 1e-6 of each
other per (baseline, product) for FP32 and FP16 and within 1e-4 for FP8
(an FP8 MFMA reduces 64 products per instruction, and moving the FFTs
between K chunks changes those groups; the FP8 visibilities' own error
against FP64 is about 4e-2), counts exact; FP32 within 1e-5 of the oracle's
restatement.  A batched call recorded as one graph launch equals its
plain launches bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

pytestmark = pytest.mark.gpu

LO, BW = 1.65e9, 32e6


def _model(nant, Ts, seed):
    # delays up to +-2 us and file offsets off by up to 1.4 samples, so the
    # first and last FFTs of the buffer meet the edge clamps
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-3e-10, 3e-10, nant),
                       rng.uniform(-2e-6, 2e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-1.4, 1.4, nant) * Ts
    return delays, fileoff


def _rel(a, b):
    live = np.linalg.norm(b, axis=-1) > 0
    return (np.linalg.norm(a[live] - b[live], axis=-1) / np.linalg.norm(b[live], axis=-1)).max()


@pytest.mark.parametrize("nchan,nant,prec", [(1024, 6, "f32"), (4096, 5, "f16"), (512, 9, "f8"), (2048, 4, "f32v"),
                                             (16384, 3, "f32"), (16384, 2, "f8")])
def test_batched_fstage_equals_per_subint_calls(nchan, nant, prec):
    F = 2 * nchan
    p = orc.plan(nchan, LO, BW)
    nf, nsub = 13, 4                          # 13 FFTs per subint, 4 subints in the buffer
    delays, fileoff = _model(nant, p.Ts, 31 + nchan)
    data = synth.make_data(nant, (nf * nsub + 1) * F, 17 + nchan)
    res = {}
    taps = {}
    for mode in ("per_subint", "batched"):
        e = FxEngine(nant, nchan, LO, BW, prec=prec, max_ffts=nsub * 64, max_bytes_per_ant=data.shape[1])
        try:
            e.set_model(delays, fileoff)
            e.upload(data)
            e.enable_tap(True)
            e.reset()
            g = e.granule
            if mode == "per_subint":
                # the tap holds the FFTs of the last F-stage call, from slot 0
                slots = (nf + g - 1) // g * g
                parts = []
                for s in range(nsub):
                    e.fstage((s * nf + np.arange(nf)).astype(np.int32), nsub, s * slots)
                    parts.append(e.read_tap(nf))
                ntot = nsub * slots
                taps[mode] = np.concatenate(parts)
            else:
                ntot = (nsub * nf + g - 1) // g * g
                e.fstage(np.arange(nsub * nf, dtype=np.int32), 1, 0)
                taps[mode] = e.read_tap(nsub * nf)
            e.xstage(0, ntot)
            e.sync()
            res[mode] = e.read_vis()
        finally:
            e.close()
    assert np.array_equal(taps["per_subint"], taps["batched"])
    (ca, cca, aa, aca), (cb, ccb, ab, acb) = res["per_subint"], res["batched"]
    assert np.array_equal(cca, ccb) and np.array_equal(aca, acb)
    bar = 1e-4 if prec == "f8" else 1e-6
    assert _rel(cb, ca) <= bar and _rel(ab, aa) <= bar
    if prec in ("f32", "f32v"):
        ref = orc.process_f32(p, data, delays, fileoff, np.arange(nsub * nf, dtype=np.int32), split=1)
        assert np.array_equal(ccb, ref["count"]) and np.array_equal(acb, ref["autocount"])
        assert _rel(cb, ref["vis"]) <= 1e-5


@pytest.mark.parametrize("prec", ["f16", "f32"])
def test_batched_subint_graph_equals_plain_launches(prec):
    nchan, nant = 1024, 5
    F = 2 * nchan
    p = orc.plan(nchan, LO, BW)
    nf, nsub, nb = 11, 8, 4                   # blocks of 4 subints per launch, two blocks per X launch
    delays, fileoff = _model(nant, p.Ts, 77)
    data = synth.make_data(nant, (nf * nsub + 1) * F, 91)
    engines = []
    for graph in (True, False):
        e = FxEngine(nant, nchan, LO, BW, prec=prec, max_ffts=nsub * 64, max_bytes_per_ant=data.shape[1])
        e.set_graph(graph)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        engines.append(e)
    try:
        g = engines[0].granule
        slots = (nb * nf + g - 1) // g * g
        nblk = nsub // nb
        for rep in range(3):                  # the recorded graph serves every replay
            for b in range(nblk):
                it = (b * nb * nf + np.arange(nb * nf)).astype(np.int32)
                xs = nblk * slots if b == nblk - 1 else 0
                for e in engines:
                    assert e.subint(it, split=nblk, fft0=b * slots, xslots=xs) == slots
            ra, rb = (e.read_vis() for e in engines)
            for x, y in zip(ra, rb):
                assert np.array_equal(x, y), rep
        if prec == "f32":
            ref = orc.process_f32(p, data, delays, fileoff, np.arange(nsub * nf, dtype=np.int32), split=1)
            cross, cc, _, _ = engines[1].read_vis()
            assert np.array_equal(cc, 3 * ref["count"])
            assert _rel(cross / 3, ref["vis"]) <= 1e-5
    finally:
        for e in engines:
            e.close()


def test_batched_fstage_at_the_benchmark_shape():
    """64 antennas x 4096 channels, FP16, radix plan 2 and X tile 4: one
    F-stage call over 16 subints of 156 FFTs against one call per subint,
    both into one X-engine launch (random 2-bit input, delays within the
    buffer's edge windows)."""
    from difx_tests_amd import workloads
    w = workloads.C3
    nf, nsub = w.nffts, 16
    delays, fileoff = workloads.delay_model(w, 5)
    data = synth.random_bytes(w.nant, (nf * nsub + 1) * w.F // 2, 11)
    res = {}
    for mode in ("per_subint", "batched"):
        e = FxEngine(w.nant, w.nchan, w.lo, w.bw, prec="f16", max_ffts=nsub * (nf + 64),
                     max_bytes_per_ant=data.shape[1])
        try:
            e.set_model(delays, fileoff)
            e.upload(data)
            e.set_plan(2)
            e.set_xtile(4)
            e.reset()
            g = e.granule
            if mode == "per_subint":
                slots = (nf + g - 1) // g * g
                for s in range(nsub):
                    e.fstage((s * nf + np.arange(nf)).astype(np.int32), nsub, s * slots)
                ntot = nsub * slots
            else:
                ntot = (nsub * nf + g - 1) // g * g
                e.fstage(np.arange(nsub * nf, dtype=np.int32), 1, 0)
            e.xstage(0, ntot)
            e.sync()
            res[mode] = e.read_vis()
        finally:
            e.close()
    (ca, cca, aa, aca), (cb, ccb, ab, acb) = res["per_subint"], res["batched"]
    assert np.array_equal(cca, ccb) and np.array_equal(aca, acb)
    assert cca.min() > 0
    assert _rel(cb, ca) <= 1e-6 and _rel(ab, aa) <= 1e-6
