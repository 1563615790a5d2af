"""Data-driven FP8 scales (fx_update_fp8_scales / FX_OPT_FP8_AUTOSCALE).

Data that the analytic default scale (flat 2-bit spectrum, one rms per
component -> 16) does not suit: 4 antennas x 4096 channels with a sloped
bandpass (x[t] + 0.9 x[t-1] before 2-bit quantisation) and a strong common
tone 0.37 channel above channel 1229, whose spectral components reach
several times e4m3's largest value 448 at the default scale; zero delays, a
correlated common signal.  Two subints of 64 FFTs in one buffer.  (A tone on
a channel centre has the same phase in every FFT, so its e4m3 rounding error
is the same in every FFT and does not average: about 4% in its auto power
whatever the scale.)

* fixed analytic scales: the second subint's tone channel saturates, its
  visibilities are off by tens of percent against the FP64 oracle;
* autoscale: the first subint's autocorrelations set per (channel, signal)
  scales, the integration is reset, and the second subint's FP8 visibilities
  meet the FP8 bars against the FP64 oracle (per (baseline, product) <= 0.1,
  parallel-hand autos <= 0.02) and the tone channel's per-element error is
  below 0.1;
* fx_update_fp8_scales called by hand gives the same bits as the option.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

from test_gpu_precision import _classes

pytestmark = pytest.mark.gpu

LO, BW = 1.65e9, 64e6
NCHAN, NANT, NF = 4096, 4, 64


def _tone_data(nant, nsamp, nchan, seed, amp=1.5, c0frac=0.3, delta=0.37, slope=0.9):
    rng = np.random.default_rng(seed)
    common = [rng.standard_normal(nsamp + 1) for _ in range(2)]
    out = np.empty((nant, nsamp // 2), np.uint8)
    t = np.arange(nsamp)
    c0 = round(c0frac * nchan)
    f0 = (c0 + delta) / (2 * nchan)                        # cycles per sample, off the channel centre
    for j in range(nant):
        codes = []
        for p in range(2):
            x = common[p] / 3 + rng.standard_normal(nsamp + 1)
            x = x[1:] + slope * x[:-1]                       # sloped bandpass
            x = x + amp * np.sqrt(1 + slope ** 2) * np.cos(2 * np.pi * f0 * t + 0.3 * j + p)
            codes.append(synth.quantise(x.astype(np.float32), float(x.std())))
        out[j] = synth.pack_codes(codes[0], codes[1])
    return out, c0


def _tone_el(cross, ref, c0):
    """Per-element error of the crosses at channel c0, normalised by the FP64
    parallel-hand autos."""
    ra = ref["autos"]
    apow = np.stack([ra[:, 0].real, ra[:, 3].real], axis=1)[:, :, c0]     # [ant, pol]
    jj, kk = np.triu_indices(NANT, 1)
    err = 0.0
    for prod in range(4):
        l, m = prod // 2, prod % 2
        norm = np.sqrt(apow[jj, l] * apow[kk, m])
        err = max(err, float((np.abs(cross[:, prod, c0] - ref["vis"][:, prod, c0]) / norm).max()))
    return err


def _run(data, mode):
    F = 2 * NCHAN
    with FxEngine(NANT, NCHAN, LO, BW, prec="f8", max_ffts=64, max_bytes_per_ant=data.shape[1]) as e:
        e.set_model(np.zeros((NANT, 3)), np.zeros(NANT))
        e.upload(data)
        if mode == "option":
            e.set_fp8_autoscale(True)
        e.reset()
        e.process(np.arange(NF, dtype=np.int32), split=2)           # subint 1 primes the scales
        if mode == "manual":
            e.update_fp8_scales()
        e.reset()
        e.process(NF + np.arange(NF, dtype=np.int32), split=2)      # subint 2 is measured
        return e.read_vis()


def test_fp8_autoscale_tone_and_slope():
    F = 2 * NCHAN
    data, c0 = _tone_data(NANT, (2 * NF + 1) * F, NCHAN, 11)
    p = orc.plan(NCHAN, LO, BW)
    ref = orc.process_f64(p, data, np.zeros((NANT, 3)), np.zeros(NANT), NF + np.arange(NF, dtype=np.int32),
                          split=2)
    # the tone exceeds e4m3's range at the analytic scale
    sigma = np.sqrt(F * 4.3128 / 2)
    peak = np.sqrt(ref["autos"][:, 0, c0].real.max() / NF) / sigma * 16
    assert peak > 448 * 1.5, peak
    fixed = _run(data, "fixed")
    auto = _run(data, "option")
    manual = _run(data, "manual")
    for x, y in zip(auto, manual):
        assert np.array_equal(x, y)
    for r in (fixed, auto):
        assert np.array_equal(r[1], ref["count"]) and np.array_equal(r[3], ref["autocount"])
    el_fixed = _tone_el(fixed[0], ref, c0)
    el_auto = _tone_el(auto[0], ref, c0)
    stmt = _classes(auto[0], auto[2], ref, NANT)
    print("FP8SCALES", {"tone_el_fixed": el_fixed, "tone_el_auto": el_auto, "auto": stmt}, flush=True)
    assert el_fixed > 0.2, el_fixed
    assert el_auto <= 0.1, el_auto
    for c, v in stmt.items():
        assert v["bp_max"] <= (0.02 if c == "auto-par" else 0.1), (c, v)
