"""One subint as one hipGraph launch (fx_subint; fx_process and the kernel.h
entry point use it).

Every replay is checked against plain launches of the same kernels (the
engine with FX_OPT_GRAPH 0) bit for bit, with the engines' own eager work
and reads between the replays: the delay tables, FFT indices and buffer
origins change from subint to subint while the recorded graphs stay the same
(they read them from device memory).  Batches of subints put the X-engine on
the last subint's graph only.  The streamed path (two ingest slots) likewise,
and everything against the oracle's FP32 restatement at 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

pytestmark = pytest.mark.gpu

LO, BW = 1.65e9, 32e6


def _model(nant, Ts, seed):
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-16, 1e-16, nant), rng.uniform(-3e-10, 3e-10, nant),
                       rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-1.4, 1.4, nant) * Ts
    return delays, fileoff


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("nchan,nant,prec", [(1024, 4, "f32"), (4096, 6, "f16"), (512, 9, "f8"), (2048, 3, "f32v")])
def test_subint_graph_equals_plain_launches(nchan, nant, prec):
    F = 2 * nchan
    p = orc.plan(nchan, LO, BW)
    nf, nsub, kb = 13, 6, 3                   # 13 FFTs per subint, 3 subints per X launch
    delays, fileoff = _model(nant, p.Ts, nchan)
    data = synth.make_data(nant, (nf * nsub + 1) * F, 5 + nchan)
    engines = []
    for graph in (True, False):
        e = FxEngine(nant, nchan, LO, BW, prec=prec, max_ffts=kb * 64, max_bytes_per_ant=data.shape[1])
        e.set_graph(graph)
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        engines.append(e)
    try:
        slots = (nf + engines[0].granule - 1) // engines[0].granule * engines[0].granule
        for s in range(nsub):
            it = (s * nf + np.arange(nf)).astype(np.int32)
            b = s % kb
            xs = kb * slots if b == kb - 1 else 0
            for e in engines:
                assert e.subint(it, split=nsub, fft0=b * slots, xslots=xs) == slots
            # eager work and reads between the replays
            outs = [e.read_vis() for e in engines]
            _same(outs[0], outs[1])
            tabs = [e.read_delay_table(nf) for e in engines]
            assert np.array_equal(tabs[0], tabs[1])
        ms, n = engines[0].graph_timing()
        cross, cc, autos, ac = engines[0].read_vis()
        if prec in ("f32", "f32v"):
            ref = orc.process_f32(p, data, delays, fileoff, np.arange(nf * nsub), split=1)
            assert np.array_equal(cc, ref["count"])
            live = cc > 0
            err = np.linalg.norm(cross[live] - ref["vis"][live], axis=-1) / np.linalg.norm(ref["vis"][live], axis=-1)
            assert err.max() <= 1e-5
    finally:
        for e in engines:
            e.close()


def test_subint_graph_streaming_slots():
    nchan, nant, nf, nsub = 1024, 5, 16, 5
    F = 2 * nchan
    p0 = orc.plan(nchan, LO, BW)
    delays, fileoff = _model(nant, p0.Ts, 77)
    stream = synth.make_data(nant, (nsub * nf + 1) * F, 78)
    nbytes = (nf + 1) * F // 2
    res = []
    for graph in (True, False):
        with FxEngine(nant, nchan, LO, BW, prec="f32", max_ffts=nf, max_bytes_per_ant=nbytes) as e:
            e.set_graph(graph)
            e.set_model(delays, fileoff)
            e.reset()
            for s in range(nsub):
                k = s % 2
                e.ingest_wait(k)
                e.ingest_host(k)[:, :nbytes] = stream[:, s * nf * F // 2: s * nf * F // 2 + nbytes]
                e.ingest_submit(k, nbytes)
                it = (s * nf + np.arange(nf)).astype(np.int32)
                ns = e.subint(it, slot=k, origin=s * nf, xslots=nf)
                assert ns == nf
            e.sync()
            res.append(e.read_vis())
    _same(res[0], res[1])
    vis = cnt = 0
    for s in range(nsub):
        p = orc.plan(nchan, LO, BW, origin=s * nf)
        buf = np.ascontiguousarray(stream[:, s * nf * F // 2: s * nf * F // 2 + nbytes])
        r = orc.process_f32(p, buf, delays, fileoff, (s * nf + np.arange(nf)).astype(np.int32))
        vis = vis + r["vis"].astype(np.complex128)
        cnt = cnt + r["count"]
    cross, cc = res[0][0], res[0][1]
    assert np.array_equal(cc, cnt)
    live = cc > 0
    err = np.linalg.norm(cross[live] - vis[live], axis=-1) / np.linalg.norm(vis[live], axis=-1)
    assert err.max() <= 1e-5


def test_subint_rejects_bad_xslots():
    with FxEngine(2, 512, LO, BW, prec="f32", max_ffts=16) as e:
        it = np.arange(8, dtype=np.int32)
        for bad in (-1, 3, 8, 32):
            with pytest.raises(Exception):
                e.subint(it, fft0=8, xslots=bad)
