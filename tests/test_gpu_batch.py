"""Batches of subbands through the C ABI (fx_batch_*): one hipGraph launch per
subint versus plain launches, against the oracle and against one engine per
subband run on its own.

Bars: FP32 subbands within 1e-5 of the oracle per (baseline, product);
graph, launch and single-engine results identical bit for bit (the same
kernels in the same order per subband); every graph replay checked against
the eager result with reads of the visibilities in between.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxBatch, FxEngine, synth

pytestmark = pytest.mark.gpu

LO = 1.65e9
# (nchan, nant, bw, prec, nffts)
SUBBANDS = [(256, 5, 8e6, "f32", 12), (1024, 3, 32e6, "f16", 6), (4096, 4, 64e6, "f32", 3),
            (16384, 3, 128e6, "f8", 2), (512, 6, 16e6, "f32v", 8)]


def _subband(nchan, nant, bw, nf, seed):
    p = orc.plan(nchan, LO, bw)
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-3e-11, 3e-11, nant),
                       rng.uniform(-2e-6, 2e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-1.3, 1.3, nant) * p.Ts
    data = synth.make_data(nant, (3 * nf + 1) * 2 * nchan, seed)
    return p, data, delays, fileoff


def _engines(cases):
    engs = []
    for (nchan, nant, bw, prec, nf), (p, data, delays, fileoff) in cases:
        e = FxEngine(nant, nchan, LO, bw, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1])
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        engs.append(e)
    return engs


def _cases(seed0=300):
    return [(sb, _subband(sb[0], sb[1], sb[2], sb[4], seed0 + i)) for i, sb in enumerate(SUBBANDS)]


def test_batch_graph_matches_launches_and_single_engines():
    cases = _cases()
    iters = [np.arange(sb[4], dtype=np.int32) + sb[4] for sb, _ in cases]    # second subint of three
    results = {}
    for mode in ("graph", "launch", "single"):
        engs = _engines(cases)
        if mode == "single":
            for e, it in zip(engs, iters):
                e.process(it, split=3)
        else:
            with FxBatch(engs) as b:
                b.process(iters, split=3, mode=mode)
                b.sync()
        results[mode] = [e.read_vis() for e in engs]
        for e in engs:
            e.close()
    for mode in ("graph", "launch"):
        for a, b in zip(results[mode], results["single"]):
            for x, y in zip(a, b):
                assert np.array_equal(x, y), mode
    # FP32 subbands against the oracle
    for ((nchan, nant, bw, prec, nf), (p, data, delays, fileoff)), res, it in zip(cases, results["graph"], iters):
        if prec not in ("f32", "f32v"):
            continue
        ref = orc.process_f32(p, data, delays, fileoff, it, split=3)
        cross, cc, autos, ac = res
        assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
        live = cc > 0
        err = np.linalg.norm(cross[live] - ref["vis"][live], axis=-1) / np.linalg.norm(ref["vis"][live], axis=-1)
        assert err.max() <= 1e-5


def test_graph_replays_against_eager():
    cases = _cases(400)
    engs_g = _engines(cases)
    engs_e = _engines(cases)
    with FxBatch(engs_g) as b:
        for sub in range(3):
            iters = [np.arange(sb[4], dtype=np.int32) + sub * sb[4] for sb, _ in cases]
            b.process(iters, split=3, mode="graph")
            for e, it in zip(engs_e, iters):
                e.process(it, split=3)
            # eager reads and synchronisations between the replays
            for eg, ee in zip(engs_g, engs_e):
                for x, y in zip(eg.read_vis(), ee.read_vis()):
                    assert np.array_equal(x, y), sub
        assert b.captures == 1
        # a different subint length records the graph again
        iters = [np.arange(sb[4] - 1, dtype=np.int32) for sb, _ in cases]
        b.process(iters, split=3, mode="graph")
        for e, it in zip(engs_e, iters):
            e.process(it, split=3)
        for eg, ee in zip(engs_g, engs_e):
            for x, y in zip(eg.read_vis(), ee.read_vis()):
                assert np.array_equal(x, y)
        assert b.captures == 2
    for e in engs_g + engs_e:
        e.close()


# 70 antennas: 5 signal blocks, partial last panels (at tile 4 the FP16 and
# FP8 kernels that load K chunks straight into two alternating LDS stages,
# over K loops of 1 to 5 chunks: 16 FFTs per FP16 chunk, 64 per FP8 chunk);
# 48 and 64 antennas: 3 and 4 blocks, one panel pair per channel at tile 4
# (the FP16 kernel that stages one panel and keeps two K chunks in flight),
# over K loops of 1, 3, 4 and 5 chunks of 16 FFTs; 1 to 16 antennas FP32:
# one block per channel, at tile 1 formed on 16x16 MFMAs (two groups per
# MFMA at up to 8 signals, one quarter at up to 16, three above), at tiles 2
# and 4 as a 32x32 block
@pytest.mark.parametrize("prec,nant,nf", [("f32", 1, 16), ("f32", 3, 48), ("f32", 4, 16), ("f32", 4, 40),
                                          ("f32", 8, 16), ("f32", 12, 16), ("f32", 16, 48), ("f32", 16, 4),
                                          ("f32", 70, 4), ("f16", 70, 4), ("f32v", 70, 4), ("f16", 70, 48),
                                          ("f16", 70, 80), ("f8", 70, 64), ("f8", 70, 128), ("f8", 70, 192),
                                          ("f16", 48, 4), ("f16", 48, 48), ("f16", 64, 64), ("f16", 64, 80)])
def test_xtile_does_not_change_results(prec, nant, nf):
    nchan = 256
    p, data, delays, fileoff = _subband(nchan, nant, 16e6, nf, 77)
    out = {}
    for sb in (1, 2, 4):
        with FxEngine(nant, nchan, LO, 16e6, prec=prec, max_ffts=nf, max_bytes_per_ant=data.shape[1]) as e:
            e.set_xtile(sb)
            assert e.xtile() == sb
            e.set_model(delays, fileoff)
            e.upload(data)
            e.reset()
            e.process(np.arange(nf, dtype=np.int32), split=3)
            out[sb] = e.read_vis()
    # tiles 2 and 4 sum each block's K in one chain, in the same order: equal
    # bits; tile 1 splits K over the four waves of a workgroup (four partial
    # chains added at the end): FP32 rounding apart, same counts
    for x, y in zip(out[4], out[2]):
        assert np.array_equal(x, y)
    cross1, cc1, autos1, ac1 = out[1]
    cross2, cc2, autos2, ac2 = out[2]
    assert np.array_equal(cc1, cc2) and np.array_equal(ac1, ac2)
    if cross2.size:                          # (one antenna: autocorrelations only)
        assert np.abs(cross1 - cross2).max() <= 1e-6 * np.abs(cross2).max()
    assert np.abs(autos1 - autos2).max() <= 1e-6 * np.abs(autos2).max()


@pytest.mark.parametrize("nant", [100, 112])
def test_partial_diagonal_panel_fp32(nant):
    # 7 signal blocks: the last diagonal panel of a 4x4 tile holds 3 blocks
    nchan, nf = 256, 3
    p, data, delays, fileoff = _subband(nchan, nant, 16e6, nf, 90 + nant)
    it = np.arange(nf, dtype=np.int32)
    with FxEngine(nant, nchan, LO, 16e6, prec="f32", max_ffts=nf, max_bytes_per_ant=data.shape[1]) as e:
        e.set_xtile(4)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it, split=3)
        cross, cc, autos, ac = e.read_vis()
    ref = orc.process_f32(p, data, delays, fileoff, it, split=3)
    assert np.array_equal(cc, ref["count"])
    live = cc > 0
    err = np.linalg.norm(cross[live] - ref["vis"][live], axis=-1) / np.linalg.norm(ref["vis"][live], axis=-1)
    assert err.max() <= 1e-5
    alive = ac > 0
    erra = np.linalg.norm(autos[alive] - ref["autos"][alive], axis=-1) / np.linalg.norm(ref["autos"][alive], axis=-1)
    assert erra.max() <= 1e-5
