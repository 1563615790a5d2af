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


@pytest.mark.parametrize("plan", [0, 1, 2, 3])
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


def _e4m3(x):
    """Round float32 values in [-448, 448] to OCP e4m3 (3 mantissa bits,
    exponent bias 7, subnormal step 2^-9), ties to even, as float64."""
    a = np.abs(x.astype(np.float64))
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -6)))
    step = 2.0 ** (e - 3)
    return np.sign(x) * np.round(a / step) * step


@pytest.mark.parametrize("nchan", [8192, 16384])
def test_complex_fp8_half_rows_vs_emulation(nchan):
    """At F >= 8192 the complex-input FP8 F-stage stores each row of eight
    FFTs as two halves of four.  The FP8 visibilities must equal those formed
    on the CPU from the F-stage's own FP32 spectra (tap): scaled by 1/s,
    clamped to +-448, rounded to e4m3, cross-multiplied in FP64 and scaled by
    s_i s_j.  Bar: 1e-4 per (baseline, product) (the MFMA's FP32 sums); a half
    row in the wrong FFT slots or rows fails it by orders of magnitude.  The
    kernel without the tap gives the same visibilities within the same bar."""
    nant, nf = 3, 16
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 90 + nchan, correlated=True)

    def run(tap, scales=None):
        with FxEngine(nant, nchan, LO, 32e6, prec="f8", max_ffts=nf, max_bytes_per_ant=data.shape[1],
                      complex8=True) as e:
            if scales is not None:
                e.set_fp8_scales(scales)
            e.set_model(delays, fileoff)
            e.upload(data)
            e.enable_tap(tap)
            e.reset()
            e.process(it)
            return (e.read_tap(nf) if tap else None), e.read_vis()

    spec, _ = run(True)                                   # [nf, nant, 2, nchan] FP32 spectra
    # per-(channel, signal) scales around the spectra's rms mapped to 16
    rng = np.random.default_rng(91)
    rms = np.sqrt(np.mean(np.abs(spec) ** 2, axis=0) / 2.0).reshape(2 * nant, nchan).T   # [nchan, 2 nant]
    scales = (np.maximum(rms, 1.0) / 16.0 * rng.uniform(0.5, 2.0, rms.shape)).astype(np.float32)
    spec2, (cross, cc, autos, ac) = run(True, scales)
    _, (cross_nt, cc_nt, _, _) = run(False, scales)
    assert np.array_equal(spec, spec2) and np.array_equal(cc, cc_nt)
    inv = (np.float32(1.0) / scales).astype(np.float32)
    q = np.empty(spec.shape, np.complex128)
    for j in range(nant):
        for l in range(2):
            s_inv = inv[:, 2 * j + l]
            re = np.clip((spec[:, j, l].real * s_inv).astype(np.float32), -448, 448)
            im = np.clip((spec[:, j, l].imag * s_inv).astype(np.float32), -448, 448)
            q[:, j, l] = _e4m3(re) + 1j * _e4m3(im)
    valid = np.any(spec != 0, axis=(2, 3))                # [nf, nant]
    b = 0
    for j in range(nant):
        for k in range(j + 1, nant):
            both = valid[:, j] & valid[:, k]
            assert cc[b] == both.sum() > 0
            for l in range(2):
                for m in range(2):
                    v = (q[both, j, l] * np.conj(q[both, k, m])).sum(axis=0)
                    v = v * scales[:, 2 * j + l].astype(np.float64) * scales[:, 2 * k + m]
                    err = np.linalg.norm(cross[b, 2 * l + m] - v) / np.linalg.norm(v)
                    assert err <= 1e-4, (j, k, l, m, err)
                    err_nt = np.linalg.norm(cross_nt[b, 2 * l + m] - cross[b, 2 * l + m]) / np.linalg.norm(v)
                    assert err_nt <= 1e-4, (j, k, l, m, err_nt)
            b += 1
