"""Streaming subints through the C ABI: global FFT indices with a buffer
origin, two pinned input slots with H2D copies overlapping the F-stage and
X-engine, against the oracle with the same origins.

Bars: FP32 visibilities of the whole stream within 1e-5 of the oracle (sum of
per-subint oracle runs with their origins) per (baseline, product), counts
exact; streaming and plain uploads bit-identical.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth

pytestmark = pytest.mark.gpu

LO = 1.65e9


def _stream(nchan, nant, nf, S, seed):
    F = 2 * nchan
    p = orc.plan(nchan, LO, 16e6)
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-15, 1e-15, nant), rng.uniform(-4e-10, 4e-10, nant),
                       rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-1.2, 1.2, nant) * p.Ts
    stream = synth.make_data(nant, (S * nf + 1) * F, seed)
    bufs = [np.ascontiguousarray(stream[:, s * nf * F // 2: (s * nf + nf + 1) * F // 2]) for s in range(S)]
    return p, delays, fileoff, bufs


def _oracle(nchan, nant, nf, S, delays, fileoff, bufs):
    vis = cnt = autos = acnt = 0
    for s in range(S):
        ps = orc.plan(nchan, LO, 16e6, origin=s * nf)
        r = orc.process_f32(ps, bufs[s], delays, fileoff, (s * nf + np.arange(nf)).astype(np.int32))
        vis = vis + r["vis"].astype(np.complex128)
        autos = autos + r["autos"].astype(np.complex128)
        cnt = cnt + r["count"]
        acnt = acnt + r["autocount"]
    return vis, cnt, autos, acnt


@pytest.mark.parametrize("nchan,nant,nf,S", [(512, 4, 8, 5), (4096, 3, 4, 4)])
def test_streaming_matches_oracle_and_plain_uploads(nchan, nant, nf, S):
    p, delays, fileoff, bufs = _stream(nchan, nant, nf, S, 500 + nchan)
    nbytes = bufs[0].shape[1]
    with FxEngine(nant, nchan, LO, 16e6, prec="f32", max_ffts=2 * (nf + 64), max_bytes_per_ant=nbytes) as e:
        e.set_model(delays, fileoff)
        e.reset()
        hosts = [e.ingest_host(0), e.ingest_host(1)]
        gran = e.granule
        slots = (nf + gran - 1) // gran * gran
        for s in range(S):
            k = s % 2
            e.ingest_wait(k)                      # the slot's host buffer is free again
            hosts[k][:, :nbytes] = bufs[s]
            e.ingest_submit(k, nbytes)
            e.fstage_ex((s * nf + np.arange(nf)).astype(np.int32), slot=k, origin=s * nf, fft0=k * slots)
            if k == 1 or s == S - 1:
                e.xstage(0, (k + 1) * slots)      # one X launch per pair of subints
        cross, cc, autos, ac = e.read_vis()
    with FxEngine(nant, nchan, LO, 16e6, prec="f32", max_ffts=2 * (nf + 64), max_bytes_per_ant=nbytes) as e2:
        e2.set_model(delays, fileoff)
        e2.reset()
        for s in range(S):
            e2.upload(bufs[s])
            e2.fstage_ex((s * nf + np.arange(nf)).astype(np.int32), origin=s * nf)
            e2.xstage(0, slots)
        plain = e2.read_vis()
    vis, cnt, a64, acnt = _oracle(nchan, nant, nf, S, delays, fileoff, bufs)
    assert np.array_equal(cc, cnt) and np.array_equal(ac, acnt)
    live = cc > 0
    err = np.linalg.norm(cross[live] - vis[live], axis=-1) / np.linalg.norm(vis[live], axis=-1)
    assert err.max() <= 1e-5
    # the plain path accumulates per subint, the streaming path per pair of
    # subints: same products, FP32 sums in a different grouping
    assert np.array_equal(plain[1], cc)
    errp = np.linalg.norm(plain[0][live] - cross[live], axis=-1) / np.linalg.norm(cross[live], axis=-1)
    assert errp.max() <= 1e-6
