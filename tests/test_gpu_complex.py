"""Complex 8-bit input on the GPU (COMPLEX 1, NBIT 8: F = nchan, Ts = 1/bw,
every FFT bin a channel) against the oracle's complex restatement, through
the C ABI.

Bars: FP32 spectra <= 1e-5 per (FFT, antenna, pol) for every radix plan;
FP32 visibilities <= 1e-5 per (baseline, product), counts exact; FP16 <=
1e-3 and FP8 <= 0.1 (crosses) against the FP64 oracle.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

pytestmark = pytest.mark.gpu

LO = 1.65e9


def _case(nchan, nant, nf, seed, bw=32e6, correlated=False):
    p = orc.plan(nchan, LO, bw, complex8=True)
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-3e-11, 3e-11, nant),
                       rng.uniform(-2e-6, 2e-6, nant)], axis=1)
    perturb = rng.uniform(-1.4, 1.4, nant)
    perturb[0] = -1.0                                   # the o = -1 clamp at FFT 0
    fileoff = delays[:, 2] + perturb * p.Ts
    nsamp = (nf + 1) * nchan
    data = (synth.make_data_complex8(nant, nsamp, seed) if correlated
            else synth.random_bytes_complex8(nant, nsamp, seed))
    return p, data, delays, fileoff, np.arange(nf, dtype=np.int32)


def _rel(a, b, axis=-1):
    return np.linalg.norm(a - b, axis=axis) / np.maximum(np.linalg.norm(b, axis=axis), 1e-30)


@pytest.mark.parametrize("plan", [0, 1, 2])
@pytest.mark.parametrize("nchan", [512, 1024, 2048, 4096, 8192, 16384])
def test_complex_fstage_fp32(nchan, plan):
    nant, nf = 3, 4
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 60 + nchan)
    with FxEngine(nant, nchan, LO, 32e6, prec="f32", max_ffts=nf, max_bytes_per_ant=data.shape[1],
                  complex8=True) as e:
        if plan >= e.nplans():
            pytest.skip("one radix plan at this length")
        assert e.F == nchan
        e.set_plan(plan)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.enable_tap(True)
        e.fstage(it)
        chan = e.read_tap(nf)
    ref = orc.process_f32(p, data, delays, fileoff, it, want_chan=True)
    v = ref["valid"].astype(bool)
    assert _rel(chan[v], ref["chan"][v]).max() <= 1e-5
    assert not np.any(chan[~v])


@pytest.mark.parametrize("prec", ["f32", "f32v"])
@pytest.mark.parametrize("nchan,nant,nf", [(512, 5, 8), (4096, 4, 4), (16384, 3, 2)])
def test_complex_vis_fp32(nchan, nant, nf, prec):
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 70 + nchan)
    with FxEngine(nant, nchan, LO, 32e6, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1],
                  complex8=True) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f32(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-5
    alive = ac > 0
    assert _rel(autos[alive], ref["autos"][alive]).max() <= 1e-5


@pytest.mark.parametrize("prec,tol", [("f16", 1e-3), ("f8", 0.1)])
def test_complex_low_precision_vs_fp64(prec, tol):
    nchan, nant, nf = 1024, 4, 32
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 80, correlated=True)
    with FxEngine(nant, nchan, LO, 32e6, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1],
                  complex8=True) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f64(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= tol
