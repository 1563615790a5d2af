"""Long integrations.

* Global FFT indices late in a second (15,600 at 64 antennas' geometry:
  4096 channels, 128 Msps; 31,200 at 1024 channels, 64 Msps), delay rates up
  to 3 us/s, real 2-bit and complex 8-bit input, through a buffer whose first
  FFT is the subint's (fx_fstage_ex origin): the FP64 delay table (a, b,
  fractional delay, offset, validity; kernel.c:373-386, 736-756 at FFT
  iteration[k], kernel.c:733) bit-exact, FP32 spectra and visibilities within
  1e-5 of the oracle's FP32 restatement.
* One full 1 s integration at 4096 channels, 128 Msps (100 subints of 156
  FFTs, 15,600 FFTs, global indices 0 ... 15,599 and maxoffset of the whole
  second), 4 antennas, delay rates up to 3 us/s, correlated 2-bit data: the
  FP32, FP16 and FP8 paths against the FP64 oracle, per class, printed as a
  LONGINT JSON line, with the bars of the short-integration statement.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth, workloads

from test_gpu_precision import _classes

pytestmark = pytest.mark.gpu


def _rel(a, b, axis=-1):
    return np.linalg.norm(a - b, axis=axis) / np.maximum(np.linalg.norm(b, axis=axis), 1e-30)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("wname,g0", [("c3-64ant-4096ch", 15600), ("c5-256ant-1024ch", 31200)])
def test_late_fft_indices(wname, g0, cplx):
    nant, nf = 6, 16
    w = workloads.WORKLOADS[wname].with_(nant=nant, rate_max=3e-6)
    F = w.nchan if cplx else w.F
    Ts = 1.0 / w.bw if cplx else w.Ts
    delays, _ = workloads.delay_model(w, 3100 + g0)
    assert np.abs(delays[:, 1]).max() / (w.F / w.fs) > 1e-6      # delay rates of order us/s
    # file offsets that put each net delay within a few samples of zero in the
    # middle of this buffer, spread so that both edge clamps and odd offsets occur
    gm = g0 + nf / 2
    tau = delays[:, 0] * gm * gm + delays[:, 1] * gm + delays[:, 2]
    fileoff = np.round(tau / Ts) * Ts + np.linspace(-2.4, 2.4, nant) * Ts
    it = (g0 + np.arange(nf)).astype(np.int32)
    if cplx:
        data = synth.make_data_complex8(nant, (nf + 1) * F, g0)
    else:
        data = synth.make_data(nant, (nf + 1) * F, g0)
    p = orc.plan(w.nchan, w.lo, w.bw, origin=g0, complex8=cplx)
    ref = orc.process_f32(p, data, delays, fileoff, it, want_chan=True)
    with FxEngine(nant, w.nchan, w.lo, w.bw, prec="f32", max_ffts=nf, max_bytes_per_ant=data.shape[1],
                  complex8=cplx) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.enable_tap(True)
        e.reset()
        ns = e.fstage_ex(it, origin=g0)
        e.xstage(0, ns)
        tab = e.read_delay_table(nf)
        spec = e.read_tap(nf)
        cross, cc, autos, ac = e.read_vis()
    assert np.array_equal(tab["valid"], ref["valid"])
    assert np.array_equal(tab["offset"], ref["offset"])
    assert np.array_equal(tab["fd"], ref["fracdelay"])
    for k in range(nf):
        for j in range(nant):
            m, a, b = orc.station_delay(delays[j], int(it[k]))
            assert tab["a"][k, j] == a and tab["b"][k, j] == b
    assert 0 < ref["valid"].sum() < nf * nant                   # edges exercised
    assert len(set(ref["offset"][ref["valid"] > 0] & 1)) == 2   # odd and even offsets
    v = ref["valid"] > 0
    err = _rel(spec[v], ref["chan"][v])
    assert err.max() <= 1e-5, err.max()
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-5


def test_one_second_integration_vs_fp64():
    nant = 4
    w = workloads.C3.with_(nant=nant)
    nf, nsub = w.nffts, 100                                     # 156 FFTs per 10 ms subint
    ntot = nf * nsub
    delays, fileoff = workloads.delay_model(w, 20261204)
    assert np.abs(delays[:, 1]).max() / (w.F / w.fs) > 1e-6
    # correlated stream: each subint's common signal at the model's integer
    # net delay in its middle (one more chunk holds the buffer's last FFT)
    shifts = np.stack([np.round(workloads.net_delay_samples(w, delays, fileoff, s * nf + nf / 2)).astype(np.int64)
                       for s in range(nsub + 1)])
    data = synth.make_stream(nant, nsub + 1, nf * w.F, 20261204, shifts)[:, : (ntot + 1) * w.F // 2]
    data = np.ascontiguousarray(data)
    p = orc.plan(w.nchan, w.lo, w.bw)
    ref = orc.process_f64(p, data, delays, fileoff, np.arange(ntot, dtype=np.int32))
    assert ref["count"].min() > 0.99 * ntot
    stmt = {"workload": f"{nant} antennas x {w.nchan} channels, 128 Msps, 1 s", "ffts": ntot, "subints": nsub,
            "reference": "FP64 oracle"}
    for prec in ("f32", "f16", "f8"):
        with FxEngine(nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1]) as e:
            e.set_model(delays, fileoff)
            e.upload(data)
            e.reset()
            for s in range(nsub):
                ns = e.fstage((s * nf + np.arange(nf)).astype(np.int32), split=nsub)
                e.xstage(0, ns)
            cross, cc, autos, ac = e.read_vis()
        assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
        stmt[prec] = _classes(cross, autos, ref, nant)
    print("LONGINT " + json.dumps(stmt), flush=True)
    for c, v in stmt["f32"].items():
        assert v["bp_max"] <= 1e-5, (c, v)
    for c, v in stmt["f16"].items():
        assert v["bp_max"] <= 1e-3, (c, v)
    for c, v in stmt["f8"].items():
        assert v["bp_max"] <= (0.02 if c == "auto-par" else 0.1), (c, v)
