"""The scaled-up FP32 workload and the uneven-subband batch against the oracle.

* 16 antennas x 2048 channels (136 baselines with autos), 128 Msps, fringe
  rates up to +-50 Hz and fractional delays U[0,1) sample: FP32 spectra on the
  MFMA X-engine and on the vector-ALU X-engine, 32 FFTs, every (baseline,
  product) within 1e-5 of the oracle's FP32 restatement of
  processAntennasAndBaselinePara (kernel.c:709-854), counts exact.
* Eight uneven subbands of a 16-antenna array (256 ... 16384 channels, 16-256
  Msps, 8-16 antennas each) through one batch: the recorded graph gives the
  same bits as plain launches at the full 10 ms FFT counts, and at a reduced
  FFT count every subband's visibilities are within 1e-5 of the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxBatch, FxEngine, synth, workloads

pytestmark = pytest.mark.gpu


def _bp_err(got, want):
    return np.linalg.norm(got - want, axis=-1) / np.linalg.norm(want, axis=-1)


@pytest.mark.parametrize("prec", ["f32", "f32v"])
def test_c2_16ant_2048ch_fringe_fracdelay(prec):
    w = workloads.C2
    nf = 32
    data, delays, fileoff, it = workloads.make_subint(w, 20261204, nffts=nf)
    # the model really has fringe rates and fractional delays
    assert np.abs(delays[:, 1]).max() * w.lo / (w.F / w.fs) > 1.0          # > 1 Hz
    frac = ((delays[:, 2] - fileoff) / w.Ts) % 1.0
    assert frac.std() > 0.1
    p = orc.plan(w.nchan, w.lo, w.bw)
    ref = orc.process_f32(p, data, delays, fileoff, it)
    with FxEngine(w.nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1]) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    assert np.array_equal(cc, ref["count"])
    assert np.array_equal(ac, ref["autocount"])
    assert (cc > 0).all()
    err = _bp_err(cross, ref["vis"])
    assert err.max() <= 1e-5, err.max()
    aerr = _bp_err(autos, ref["autos"])
    assert aerr.max() <= 1e-5, aerr.max()


def _subband_engines(nffts_of, seed, correlated=True, prec="f32"):
    engines, iters, cases = [], [], []
    for i, w in enumerate(workloads.C4):
        nf = nffts_of(w)
        if correlated:
            data, delays, fileoff, it = workloads.make_subint(w.with_(rate_max=1e-9), seed + i, nffts=nf)
        else:
            delays, fileoff = workloads.delay_model(w.with_(rate_max=1e-9), seed + i)
            data = synth.random_bytes(w.nant, (nf + 1) * w.F // 2, seed + i)
            it = np.arange(nf, dtype=np.int32)
        e = FxEngine(w.nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1])
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        engines.append(e)
        iters.append(it)
        cases.append((w, data, delays, fileoff, it))
    return engines, iters, cases


def test_c4_subbands_graph_equals_launches_full_counts():
    engines, iters, cases = _subband_engines(lambda w: w.nffts, 4100, correlated=False)
    try:
        assert sorted({w.nchan for w, *_ in cases}) == [256, 512, 1024, 2048, 4096, 8192, 16384]
        with FxBatch(engines) as b:
            out = {}
            for mode in ("launch", "graph", "graph"):
                for e in engines:
                    e.reset()
                b.process(iters, 1, mode)
                b.sync()
                out.setdefault(mode, []).append([e.read_vis() for e in engines])
            assert b.captures == 1
        for runs in out.values():
            for r in runs:
                for (x, xc, a, acnt), (y, yc, bb, bc) in zip(r, out["launch"][0]):
                    assert np.array_equal(x, y) and np.array_equal(xc, yc)
                    assert np.array_equal(a, bb) and np.array_equal(acnt, bc)
    finally:
        for e in engines:
            e.close()


def test_c4_subbands_against_oracle():
    engines, iters, cases = _subband_engines(lambda w: min(w.nffts, 24), 4200)
    try:
        with FxBatch(engines) as b:
            b.process(iters, 1, "graph")
            b.sync()
        for e, (w, data, delays, fileoff, it) in zip(engines, cases):
            cross, cc, autos, ac = e.read_vis()
            ref = orc.process_f32(orc.plan(w.nchan, w.lo, w.bw), data, delays, fileoff, it)
            assert np.array_equal(cc, ref["count"]), w.name
            live = cc > 0
            assert live.all(), w.name
            err = _bp_err(cross, ref["vis"])
            assert err.max() <= 1e-5, (w.name, err.max())
    finally:
        for e in engines:
            e.close()
