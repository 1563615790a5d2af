"""GPU parity of the HIP engine against the CPU oracle, through the C ABI.

Bars (stated per check):
  * delay table: offsets and validity bit-exact; a, b, fd bit-exact (FP64,
    same operation order, no contraction on either side)
  * FP32 spectra: per (FFT, antenna, pol) ||dY|| / ||Y|| <= 1e-5 against the
    FP32 restatement of the reference's IPP path
  * FP32 visibilities: per (baseline, product) ||dV|| / ||V|| <= 1e-5, counts
    exact, autocorrelations likewise
  * FP16 visibilities: per (baseline, product) ||dV|| / ||V64|| <= 1e-3
    against the FP64 oracle
  * FP8 visibilities: per (baseline, product) ||dV|| / ||V64|| <= 0.1 for
    crosses and cross-hand autocorrelations, <= 0.02 for parallel-hand
    autocorrelations, against the FP64 oracle.  e4m3 keeps 3 mantissa bits,
    so each product carries ~5% rms rounding error; on noise-dominated
    products that error does not average down relative to the visibility.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

pytestmark = pytest.mark.gpu

LO = 1.65e9


def _case(nchan, nant, nf, seed, bw=32e6, rate=3e-11, spread=1.4, data="random"):
    """Random model with net delays of up to +-spread samples (odd offsets,
    edge clamps and invalid antennas all occur)."""
    p = orc.plan(nchan, LO, bw)
    F = 2 * nchan
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-rate, rate, nant),
                       rng.uniform(-2e-6, 2e-6, nant)], axis=1)
    perturb = rng.uniform(-spread, spread, nant)
    if spread >= 2.0:
        # net delays that hit the edge cases: invalid at the first FFT, invalid
        # at the last FFT, the o = -1 and o = maxoffset + 1 clamps (kernel.c:742-755)
        edge = [-2.0, 2.6, -1.0, 1.6, 0.49]
        perturb[: min(nant, len(edge))] = edge[: min(nant, len(edge))]
    # net delay at FFT 0 is -perturb samples (plus the small rate term)
    fileoff = delays[:, 2] + perturb * p.Ts
    if data == "random":
        d = synth.random_bytes(nant, (nf + 1) * F // 2, seed)
    else:
        d = synth.make_data(nant, (nf + 1) * F, seed)
    return p, d, delays, fileoff, np.arange(nf, dtype=np.int32)


def _engine(p, nant, nf, prec="f32", data=None):
    e = FxEngine(nant, p.nchan, LO, p.bw, prec=prec, max_ffts=nf,
                 max_bytes_per_ant=0 if data is None else data.shape[1])
    return e


def _rel(a, b, axis=-1):
    return np.linalg.norm(a - b, axis=axis) / np.maximum(np.linalg.norm(b, axis=axis), 1e-30)


@pytest.mark.parametrize("nchan", [256, 512, 1024, 2048, 4096, 8192, 16384])
def test_delay_table_bit_exact(nchan):
    nant, nf = 5, 12
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 10 + nchan, rate=2e-10, spread=2.2)
    with _engine(p, nant, nf, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.fstage(it)
        tab = e.read_delay_table(nf)
    ref = orc.process_f32(p, data, delays, fileoff, it)
    assert np.array_equal(tab["valid"], ref["valid"])
    assert np.array_equal(tab["offset"], ref["offset"])
    assert np.array_equal(tab["fd"], ref["fracdelay"])
    for k in range(nf):
        for j in range(nant):
            m, a, b = orc.station_delay(delays[j], int(it[k]))
            assert tab["a"][k, j] == a and tab["b"][k, j] == b
    assert 0 < ref["valid"].sum() < nf * nant          # the case exercises invalid antennas


@pytest.mark.parametrize("plan", [0, 1, 2, 3])
@pytest.mark.parametrize("nchan", [256, 512, 1024, 2048, 4096, 8192, 16384])
def test_fstage_fp32_matches_oracle(nchan, plan):
    nant, nf = 3, 6
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 20 + nchan)
    with _engine(p, nant, nf, data=data) as e:
        if plan >= e.nplans():
            pytest.skip(f"one radix plan at nchan={nchan}")
        e.set_plan(plan)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.enable_tap(True)
        e.fstage(it)
        chan = e.read_tap(nf)
    ref = orc.process_f32(p, data, delays, fileoff, it, want_chan=True)
    v = ref["valid"].astype(bool)
    err = _rel(chan[v], ref["chan"][v])
    assert err.max() <= 1e-5, err.max()
    assert not np.any(chan[~v])                          # invalid slots are zero


@pytest.mark.parametrize("prec", ["f32", "f32v"])
@pytest.mark.parametrize("nchan,nant,nf", [(256, 4, 16), (1024, 4, 9), (2048, 6, 5), (4096, 3, 4), (512, 17, 6),
                                           (256, 40, 8), (8192, 3, 3), (16384, 3, 3)])
def test_vis_fp32_matches_oracle(nchan, nant, nf, prec):
    # f32: FP32 MFMA X-engine; f32v: plain FP32 vector X-engine
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 30 + nchan + nant)
    with _engine(p, nant, nf, prec=prec, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f32(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = ref["count"] > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-5
    assert not np.any(cross[~live])
    alive = ref["autocount"] > 0
    assert _rel(autos[alive], ref["autos"][alive]).max() <= 1e-5


@pytest.mark.parametrize("order", ["reversed", "every-other"])
def test_vis_noncontiguous_iteration(order):
    # FFT index lists that are not a contiguous run go to the device as an
    # array (a contiguous run travels as kernel arguments): same visibilities
    # and counts as the oracle on the same list
    nchan, nant, nf = 1024, 5, 24
    p, data, delays, fileoff, _ = _case(nchan, nant, 2 * nf, 77, spread=2.5)
    it = np.arange(nf, dtype=np.int32)[::-1].copy() if order == "reversed" else np.arange(0, 2 * nf, 2, dtype=np.int32)
    with _engine(p, nant, nf, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it, split=2)                     # maxoffset covers FFT 2*nf - 2
        cross, cc, autos, ac = e.read_vis()
        table = e.read_delay_table(nf)
    ref = orc.process_f32(p, data, delays, fileoff, it, split=2)
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-5
    alive = ac > 0
    assert _rel(autos[alive], ref["autos"][alive]).max() <= 1e-5
    assert table["valid"].shape == (nf, nant) and table["valid"].sum() > nf


def test_vis_fp32_split_and_accumulation():
    # two calls with split = 2 (kernel.c:730 maxoffset), accumulated on the device
    nchan, nant, nf = 512, 4, 8
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 77, spread=1.9)
    half = nf // 2
    ref = orc.process_f32(p, data, delays, fileoff, it[:half], split=2)
    ref2 = orc.process_f32(p, data, delays, fileoff, it[half:], split=2)
    with _engine(p, nant, nf, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it[:half], split=2)
        e.process(it[half:], split=2)
        cross, cc, autos, ac = e.read_vis()
    assert np.array_equal(cc, ref["count"] + ref2["count"])
    want = ref["vis"] + ref2["vis"]
    live = cc > 0
    assert _rel(cross[live], want[live]).max() <= 1e-5


def test_x_batching_over_slots():
    # two F-stage calls into different FFT slots, one X-engine pass over both
    nchan, nant, nf = 1024, 4, 8
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 88)
    with _engine(p, nant, 2 * nf, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        n0 = e.fstage(it[:4])
        n1 = e.fstage(it[4:], fft0=n0)
        e.xstage(0, n0 + n1)
        cross, cc, autos, ac = e.read_vis()
    # each call has its own maxoffset: (4 - 1) * F
    r0 = orc.process_f32(p, data, delays, fileoff, it[:4])
    r1 = orc.process_f32(p, data, delays, fileoff, it[4:])
    assert np.array_equal(cc, r0["count"] + r1["count"])
    live = cc > 0
    assert _rel(cross[live], (r0["vis"] + r1["vis"])[live]).max() <= 1e-5


def test_invalid_antenna_contributes_nothing():
    nchan, nant, nf = 256, 4, 6
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 99, spread=0.3)
    fileoff[2] += 1e-3                                   # antenna 2 far outside its window
    with _engine(p, nant, nf, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f32(p, data, delays, fileoff, it)
    assert ref["valid"][:, 2].sum() == 0 and ac[2] == 0 and not np.any(autos[2])
    assert np.array_equal(cc, ref["count"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-5


@pytest.mark.parametrize("nchan,nant,nf", [(1024, 4, 32), (4096, 8, 16), (512, 20, 24), (16384, 4, 8)])
def test_vis_fp16_vs_fp64_oracle(nchan, nant, nf):
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 40 + nchan, data="correlated")
    with _engine(p, nant, nf, prec="f16", data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f64(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-3
    assert _rel(autos, ref["autos"]).max() <= 1e-3


@pytest.mark.parametrize("xtile", [1, 2, 4])
@pytest.mark.parametrize("prec,nf", [("f32", 23), ("f32", 40), ("f16", 40), ("f16", 75)])
def test_vis_multi_step_k_pipeline(prec, nf, xtile):
    # X-engine K loops of 3 to 10 chunks, odd and even (two chunks per
    # pipeline step, the last one half empty when odd), with 80 signals in
    # 3 blocks: panels that overhang the last block at tiles 2 and 4
    nchan, nant = 256, 40
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 300 + nf + xtile, data="correlated")
    with _engine(p, nant, nf, prec=prec, data=data) as e:
        e.set_xtile(xtile)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    if prec == "f32":
        ref, tol = orc.process_f32(p, data, delays, fileoff, it), 1e-5
    else:
        ref, tol = orc.process_f64(p, data, delays, fileoff, it), 1e-3
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= tol
    alive = ac > 0
    assert _rel(autos[alive], ref["autos"][alive]).max() <= tol


@pytest.mark.parametrize("prec,nf", [("f32", 40), ("f16", 80), ("f16", 100), ("f8", 192)])
@pytest.mark.parametrize("nchan", [256, 1024])
def test_vis_small_grid_k_pipeline(prec, nf, nchan):
    # one 32-signal block per channel and at most 1024 workgroups: the X-engine
    # stages several K chunks per step; chunk counts 5, 7 and 3 leave the last
    # step partly empty
    nant = 8
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 400 + nf + nchan, data="correlated")
    with _engine(p, nant, nf, prec=prec, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    if prec == "f32":
        ref, tol, tola = orc.process_f32(p, data, delays, fileoff, it), 1e-5, 1e-5
    else:
        ref = orc.process_f64(p, data, delays, fileoff, it)
        tol, tola = (1e-3, 1e-3) if prec == "f16" else (0.1, 0.1)
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= tol
    alive = ac > 0
    assert _rel(autos[alive], ref["autos"][alive]).max() <= tola


@pytest.mark.parametrize("nchan,nant,nf", [(1024, 4, 64), (2048, 12, 32), (16384, 4, 24)])
def test_vis_fp8_vs_fp64_oracle(nchan, nant, nf):
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 50 + nchan, data="correlated")
    with _engine(p, nant, nf, prec="f8", data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f64(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"])
    live = cc > 0
    ec = _rel(cross[live], ref["vis"][live]).max()
    ea = _rel(autos[:, [0, 3]], ref["autos"][:, [0, 3]]).max()
    ex = _rel(autos[:, [1, 2]], ref["autos"][:, [1, 2]]).max()
    print(f"fp8 max rel error: cross {ec:.3e} parallel-hand auto {ea:.3e} cross-hand auto {ex:.3e}")
    assert ec <= 0.1 and ex <= 0.1 and ea <= 0.02


@pytest.mark.parametrize("prec", ["f16", "f8"])
@pytest.mark.parametrize("nchan", [1024, 4096, 8192, 16384])
def test_radix_plans_agree_in_low_precision(nchan, prec):
    # every plan feeds the FP16 / FP8 X-engine: same counts, visibilities
    # within the FP32 spectrum difference
    nant, nf = 4, 16
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 60 + nchan, data="correlated")
    out = []
    for plan in range(4):
        with _engine(p, nant, nf, prec=prec, data=data) as e:
            if plan >= e.nplans():
                continue
            e.set_plan(plan)
            e.set_model(delays, fileoff)
            e.upload(data)
            e.reset()
            e.process(it)
            out.append(e.read_vis())
    live = out[0][1] > 0
    tol = 1e-3 if prec == "f16" else 0.05
    for o in out[1:]:
        assert np.array_equal(out[0][1], o[1])
        assert _rel(o[0][live], out[0][0][live]).max() <= tol


@pytest.mark.parametrize("prec", ["f16", "f8"])
@pytest.mark.parametrize("plan", [0, 1, 2, 3])
@pytest.mark.parametrize("nchan,nf", [(512, 32), (1024, 32), (2048, 24), (4096, 16), (8192, 16)])
def test_vis_lowprec_every_plan_vs_fp64_oracle(nchan, nf, plan, prec):
    # every radix plan of the FP16 / FP8 F-stage against the FP64 oracle: the
    # per-FFT state kept in registers across the FFT loop (twiddle factors,
    # FP8 scales, the cross-lane plan's twiddles) differs by plan and precision.
    # FP8's error in the autos falls as 1/sqrt(FFTs): at least 16 FFTs keep
    # it under the 2% bound (8 FFTs at 8192 channels measured 2.2%)
    nant = 5
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 500 + nchan + 7 * plan, data="correlated")
    with _engine(p, nant, nf, prec=prec, data=data) as e:
        if plan >= e.nplans():
            pytest.skip(f"{e.nplans()} radix plans at nchan={nchan}")
        e.set_plan(plan)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f64(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    tol, tola = (1e-3, 1e-3) if prec == "f16" else (0.1, 0.02)
    assert _rel(cross[live], ref["vis"][live]).max() <= tol
    assert _rel(autos[:, [0, 3]], ref["autos"][:, [0, 3]]).max() <= tola


@pytest.mark.parametrize("prec", ["f16", "f8"])
@pytest.mark.parametrize("nchan", [1024, 4096, 16384])
def test_vis_lowprec_invalid_ffts_vs_fp64_oracle(nchan, prec):
    # net delays that make whole FFTs of an antenna invalid (the first, the
    # last, clamped edges): the F-stage zeroes those slots of its register
    # chain (FP8: one half of a word, FFTs in pairs; at 16384 channels the
    # staged slots of a row) and the counts skip them
    nant, nf = 5, 32
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 600 + nchan, spread=2.4, data="correlated")
    with _engine(p, nant, nf, prec=prec, data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f64(p, data, delays, fileoff, it)
    assert 0 < ref["valid"].sum() < nf * nant          # some FFTs of some antennas are invalid
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live = cc > 0
    tol, tola = (1e-3, 1e-3) if prec == "f16" else (0.1, 0.02)
    assert _rel(cross[live], ref["vis"][live]).max() <= tol
    alive = ac > 0
    assert _rel(autos[alive][:, [0, 3]], ref["autos"][alive][:, [0, 3]]).max() <= tola


def test_autotune_and_cache(tmp_path, monkeypatch):
    nchan, nant, nf = 2048, 6, 12
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 71)
    cache = str(tmp_path / "tune.txt")
    with _engine(p, nant, 4 * nf, prec="f32", data=data) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        rep = e.autotune(nf, reps=2, cache=cache)
        assert rep["plan"] in range(e.nplans()) and rep["xtile"] in (1, 2, 4) and rep["xbatch"] in (1, 2, 4)
        assert e.plan() == rep["plan"] and e.xtile() == rep["xtile"] and e.xbatch() == rep["xbatch"]
        cross, cc, autos, ac = e.read_vis()              # the tuner leaves the visibilities zeroed
        assert not np.any(cc) and not np.any(cross) and not np.any(autos)
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f32(p, data, delays, fileoff, it)
    assert np.array_equal(cc, ref["count"])
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 1e-5
    lines = [l.split() for l in open(cache) if not l.startswith("#")]
    assert len(lines) == 1 and lines[0][1:4] == [str(nant), str(nchan), "0"]
    # a new context with FX_TUNE_CACHE set starts from the stored choice
    monkeypatch.setenv("FX_TUNE_CACHE", cache)
    with _engine(p, nant, nf, prec="f32", data=data) as e2:
        assert (e2.plan(), e2.xtile(), e2.xbatch()) == (rep["plan"], rep["xtile"], rep["xbatch"])


def test_fp8_per_channel_signal_scales():
    # non-uniform scales: the F-stage divides by s[c][sig] and the X-engine
    # multiplies by s_i s_j; a transposed or misread table breaks the bound
    nchan, nant, nf = 1024, 5, 64
    p, data, delays, fileoff, it = _case(nchan, nant, nf, 83, data="correlated")
    rng = np.random.default_rng(84)
    sigma = np.sqrt(2 * nchan * 4.3128 / 2.0)
    scales = (sigma / 16.0) * rng.uniform(0.5, 2.0, (nchan, 2 * nant)).astype(np.float32)
    with _engine(p, nant, nf, prec="f8", data=data) as e:
        e.set_fp8_scales(scales)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f64(p, data, delays, fileoff, it)
    live = cc > 0
    assert _rel(cross[live], ref["vis"][live]).max() <= 0.1
    assert _rel(autos[:, [0, 3]], ref["autos"][:, [0, 3]]).max() <= 0.02
