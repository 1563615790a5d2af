"""Array-size edges of the HIP engine against the oracle, through the C ABI.

* One antenna: no baselines (empty cross arrays), only the four
  autocorrelation products of its two polarisations.
* 300 antennas (600 signals, 19 padded 32-signal blocks, 190 block pairs per
  channel): more signals than any bench workload and a count that fills no
  X-engine tile evenly, at every tile size.

Bars as in test_gpu_parity: FP32 per (baseline, product) ||dV|| / ||V|| <=
1e-5 against the oracle's FP32 restatement; FP16 <= 1e-3 and FP8 <= 0.1
(crosses, cross-hand autos) / 0.02 (parallel-hand autos) against the FP64
oracle; counts exact.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

pytestmark = pytest.mark.gpu

LO, BW = 1.65e9, 32e6


def _case(nant, nchan, nf, seed):
    p = orc.plan(nchan, LO, BW)
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-3e-11, 3e-11, nant),
                       rng.uniform(-2e-6, 2e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-1.4, 1.4, nant) * p.Ts
    data = synth.make_data(nant, (nf + 1) * 2 * nchan, seed)
    return p, data, delays, fileoff, np.arange(nf, dtype=np.int32)


def _run(p, data, delays, fileoff, it, prec, xtile=None):
    nant = data.shape[0]
    with FxEngine(nant, p.nchan, LO, BW, prec=prec, max_ffts=len(it), max_bytes_per_ant=data.shape[1]) as e:
        if xtile is not None:
            e.set_xtile(xtile)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        return e.read_vis()


def _rel(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-30)


def _check(prec, got, p, data, delays, fileoff, it):
    cross, cc, autos, ac = got
    if prec == "f32":
        ref = orc.process_f32(p, data, delays, fileoff, it)
        tc = ta = tx = 1e-5
    else:
        ref = orc.process_f64(p, data, delays, fileoff, it)
        tc, ta, tx = (1e-3, 1e-3, 1e-3) if prec == "f16" else (0.1, 0.02, 0.1)
    assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
    live, alive = cc > 0, ac > 0
    ec = _rel(cross[live], ref["vis"][live]).max() if live.any() else 0.0
    ea = _rel(autos[alive][:, [0, 3]], ref["autos"][alive][:, [0, 3]]).max()
    ex = _rel(autos[alive][:, [1, 2]], ref["autos"][alive][:, [1, 2]]).max()
    print(f"EDGES {prec} nant={data.shape[0]} cross {ec:.2e} auto-par {ea:.2e} auto-x {ex:.2e}", flush=True)
    assert ec <= tc and ea <= ta and ex <= tx
    assert not np.any(cross[~live]) and not np.any(autos[~alive])
    return ref


@pytest.mark.parametrize("prec,nf", [("f32", 16), ("f16", 16), ("f8", 64)])
def test_single_antenna_autos_only(prec, nf):
    p, data, delays, fileoff, it = _case(1, 1024, nf, 71)
    got = _run(p, data, delays, fileoff, it, prec)
    cross, cc, autos, ac = got
    assert cross.shape == (0, 4, 1024) and cc.shape == (0,) and autos.shape == (1, 4, 1024)
    assert ac[0] > 0
    _check(prec, got, p, data, delays, fileoff, it)


@pytest.mark.parametrize("prec,nf,xtile", [("f32", 8, 1), ("f32", 8, 4), ("f16", 8, 1), ("f16", 8, 2),
                                           ("f16", 8, 4), ("f8", 48, 2), ("f8", 48, 4)])
def test_ragged_300_antennas(prec, nf, xtile):
    p, data, delays, fileoff, it = _case(300, 256, nf, 300 + nf)
    got = _run(p, data, delays, fileoff, it, prec, xtile)
    assert got[0].shape == (300 * 299 // 2, 4, 256)
    _check(prec, got, p, data, delays, fileoff, it)
