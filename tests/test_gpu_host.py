"""The kernel.h boundary driven by C callers, on the GPU.

* fxmain (this repository's host program, the behaviour of main.c) and, when
  it was built where the reference exists, the reference's own main.c linked
  against libdifxfx.so (oracle/build_ref_host.sh), each run on a generated
  config and data files; their vis.out must match the oracle's normalised
  visibilities (FP32 bar 1e-5 per baseline-product, plus the 6-decimal print).
* tests/kernelh_driver.c calls the per-(antenna, FFT) functions
  (getStationDelay, stationDelayAndOffset, processAntennas -> unpack,
  fringerotate, dofft, fracSampleCorrect, conjChannels, processBaselinePara,
  merge, processNormalize) and dumps their results.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import config as cfgmod
from difx_tests_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FXMAIN = os.path.join(ROOT, "difx_tests_amd", "lib", "fxmain")
REFMAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main")
DRIVER = os.path.join(ROOT, "tests", "_build", "kernelh_driver")
LO, BW = 1.65e9, 32e6


def _write_case(d, nant, nchan, numffts, seed, spread=1.3):
    p = orc.plan(nchan, LO, BW)
    rng = np.random.default_rng(seed)
    delays = np.stack([rng.uniform(-1e-17, 1e-17, nant), rng.uniform(-2e-11, 2e-11, nant),
                       rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-spread, spread, nant) * p.Ts
    cfg = cfgmod.Config(nchan=nchan, lo=LO, bandwidth=BW, numffts=numffts)
    data = synth.make_data(nant, (numffts + 1) * 2 * nchan, seed)
    for j in range(nant):
        fn = f"a{j}.bin"
        data[j].tofile(os.path.join(d, fn))
        cfg.antennas.append(cfgmod.AntennaEntry(f"A{j}", fn, tuple(delays[j]), float(fileoff[j])))
    assert data.shape[1] == cfg.subintbytes
    with open(os.path.join(d, "test.conf"), "w") as f:
        f.write(cfgmod.format_config(cfg))
    return p, data, delays, fileoff


def _read_visout(path, nbl, nchan):
    a = np.loadtxt(path, ndmin=2)
    assert a.shape == (nchan, 2 + nbl * 16)
    v = a[:, 2:].reshape(nchan, nbl, 4, 4)
    vis = (v[..., 0] + 1j * v[..., 1]).transpose(1, 2, 0)
    return a[:, :2], vis, v[..., 2].transpose(1, 2, 0), v[..., 3].transpose(1, 2, 0)


def _check_visout(d, p, data, delays, fileoff, numffts, split=1):
    nant = data.shape[0]
    nbl = nant * (nant - 1) // 2
    hdr, vis, amp, ph = _read_visout(os.path.join(d, "vis.out"), nbl, p.nchan)
    c = np.arange(p.nchan)
    assert np.allclose(hdr[:, 0], c) and np.allclose(hdr[:, 1], (c + 0.5) / p.nchan * p.bw / 1e6, atol=1e-6)
    ref = np.zeros((nbl, 4, p.nchan), np.complex128)
    cnt = np.zeros(nbl)
    nsplit = numffts // split
    for s in range(split):
        r = orc.process_f32(p, data, delays, fileoff, np.arange(s * nsplit, (s + 1) * nsplit), split=split)
        ref += r["vis"]
        cnt += r["count"]
    ref[cnt > 0] /= cnt[cnt > 0, None, None]
    err = np.linalg.norm(vis - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
    assert err.max() <= 1e-5 + 1e-6 * np.sqrt(p.nchan) / np.min(np.linalg.norm(ref, axis=-1))
    # amplitude column (ippsMagnitude_32fc, kernel.c:600): same norm bar
    assert (np.linalg.norm(amp - np.abs(ref), axis=-1) / np.linalg.norm(ref, axis=-1)).max() <= 1e-5 + 1e-5
    log = open(os.path.join(d, "runtime.fxkernel.log")).read()
    assert log.startswith("Run time was ") and "Start time was" in log


@pytest.mark.parametrize("split", [1, 2])
def test_fxmain_end_to_end(tmp_path, split):
    p, data, delays, fileoff = _write_case(str(tmp_path), 6, 256, 64, 11 + split)
    r = subprocess.run([FXMAIN, "./test.conf", str(split)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _check_visout(str(tmp_path), p, data, delays, fileoff, 64, split)


@pytest.mark.skipif(not os.path.exists(REFMAIN), reason="reference host program not built (reference absent)")
def test_reference_main_c_on_this_library(tmp_path):
    # main.c:324 reads ./test.conf; MAX_ANTENNAS is 20 (main.c:10)
    p, data, delays, fileoff = _write_case(str(tmp_path), 5, 1024, 32, 23)
    r = subprocess.run([REFMAIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _check_visout(str(tmp_path), p, data, delays, fileoff, 32)


@pytest.mark.parametrize("nchan", [512, 16384])
def test_kernelh_fine_grained_functions(tmp_path, nchan):
    nant, numffts, g = 3, 6, 4
    p = orc.plan(nchan, LO, BW)
    F = 2 * nchan
    rng = np.random.default_rng(5)
    delays = np.stack([rng.uniform(-1e-17, 1e-17, nant), rng.uniform(-3e-11, 3e-11, nant),
                       rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + np.array([0.6, -0.8, 0.2]) * p.Ts
    data = synth.random_bytes(nant, (numffts + 1) * F // 2, 5)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<5i", nant, nchan, data.shape[1], g, numffts))
        f.write(struct.pack("<2d", LO, BW))
        f.write(delays.astype("<f8").tobytes())
        f.write(fileoff.astype("<f8").tobytes())
        f.write(data.tobytes())
    out = tmp_path / "out.bin"
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    buf = open(out, "rb").read()
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(buf, dtype, n, pos)
        pos += a.nbytes
        return a

    nbl = nant * (nant - 1) // 2
    m, a, b = take("<f8", 3)
    assert (m, a, b) == orc.station_delay(delays[0], g)
    off, val = take("<i4", 2)
    fd, da, db = take("<f8", 3)
    v0, o0, fd0 = orc.offset(m, fileoff[0], p.Ts, g, F, (numffts - 1) * F)
    assert (off, val, fd) == (o0, v0, fd0) and (da, db) == (a, b)
    valid = take("<i4", nant)
    unp = take("<c8", nant * 2 * F).reshape(nant, 2, F)
    chan = take("<c8", nant * 2 * F).reshape(nant, 2, F)
    conj = take("<c8", nant * 2 * nchan).reshape(nant, 2, nchan)
    vis = take("<c8", nbl * 4 * nchan).reshape(nbl, 4, nchan)
    cnt = take("<i4", nbl)
    visn = take("<c8", nbl * 4 * nchan).reshape(nbl, 4, nchan)
    unp2 = take("<c8", nant * 2 * F).reshape(nant, 2, F)
    chan2 = take("<c8", nant * 2 * F).reshape(nant, 2, F)
    conj2 = take("<c8", nant * 2 * nchan).reshape(nant, 2, nchan)
    vis2 = take("<c8", nbl * 4 * nchan).reshape(nbl, 4, nchan)
    cnt2 = take("<i4", nbl)
    assert pos == len(buf)
    # processAntennasAndBaselinePara over FFTs 0..numffts-1: visibilities and
    # counts, and every antenna's buffers hold its last valid FFT
    full = orc.process_f32(p, data, delays, fileoff, np.arange(numffts), want_chan=True)
    assert np.array_equal(cnt2, full["count"])
    live = cnt2 > 0
    assert live.any()
    assert np.linalg.norm(vis2[live] - full["vis"][live]) / np.linalg.norm(full["vis"][live]) < 1e-5
    nlast = 0
    for j in range(nant):
        ks = np.nonzero(full["valid"][:, j])[0]
        if len(ks) == 0:
            assert not np.any(unp2[j]) and not np.any(chan2[j]) and not np.any(conj2[j])
            continue
        k = int(ks[-1])
        nlast += 1
        mj, aj, bj = orc.station_delay(delays[j], k)
        vj, oj, fdj = orc.offset(mj, fileoff[j], p.Ts, k, F, (numffts - 1) * F)
        assert vj and oj == full["offset"][k, j]
        l0, l1 = orc.unpack(data[j], oj, F)
        rot = orc.fringe_rotator(p, aj, bj)
        for pol, lv in enumerate((l0, l1)):
            u = lv.astype(np.complex128) * rot
            assert np.max(np.abs(unp2[j, pol] - u)) < 1e-5
            Y = np.fft.fft(unp2[j, pol].astype(np.complex128))
            assert np.linalg.norm(chan2[j, pol, nchan:] - Y[nchan:]) / np.linalg.norm(Y[nchan:]) < 1e-5
            want = full["chan"][k, j, pol]
            assert np.linalg.norm(chan2[j, pol, :nchan] - want) / np.linalg.norm(want) < 1e-5
            assert np.array_equal(conj2[j, pol], np.conj(chan2[j, pol, :nchan]))
    assert nlast > 0
    # processAntennas leaves the last FFT (numffts - 1), maxoffset (numffts-1)*F
    last = numffts - 1
    ref = orc.process_f32(p, data, delays, fileoff, np.array([last]), split=numffts, want_chan=True)
    assert np.array_equal(valid, ref["valid"][0])
    for j in range(nant):
        if not valid[j]:
            continue
        mj, aj, bj = orc.station_delay(delays[j], last)
        vj, oj, fdj = orc.offset(mj, fileoff[j], p.Ts, last, F, last * F)
        l0, l1 = orc.unpack(data[j], oj, F)
        rot = orc.fringe_rotator(p, aj, bj)
        for pol, lv in enumerate((l0, l1)):
            u = lv.astype(np.complex128) * rot
            assert np.max(np.abs(unp[j, pol] - u)) < 1e-5
            Y = np.fft.fft(u)
            assert np.linalg.norm(chan[j, pol, nchan:] - Y[nchan:]) / np.linalg.norm(Y[nchan:]) < 1e-5
            assert np.linalg.norm(chan[j, pol, :nchan] - ref["chan"][0, j, pol]) / np.linalg.norm(
                ref["chan"][0, j, pol]) < 1e-5
            assert np.array_equal(conj[j, pol], np.conj(chan[j, pol, :nchan]))
    assert np.array_equal(cnt, ref["count"])
    live = cnt > 0
    assert np.linalg.norm(vis[live] - ref["vis"][live]) / np.linalg.norm(ref["vis"][live]) < 1e-5
    want = ref["vis"][live] / cnt[live, None, None]
    assert np.linalg.norm(visn[live] - want) / np.linalg.norm(want) < 1e-5


FXBATCH = os.path.join(ROOT, "difx_tests_amd", "lib", "fxbatch")


def test_fxbatch_uneven_subbands(tmp_path):
    # three subbands (channels, bandwidth, LO, antenna subset), two subints of
    # a continuous stream per (subband, antenna) file; graph and plain
    # launches; each subband's normalised visibilities against the oracle
    from difx_tests_amd import visio
    rng = np.random.default_rng(31)
    nant, nsub = 5, 2
    delays = np.stack([np.zeros(nant), rng.uniform(-2e-11, 2e-11, nant), rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    spread = rng.uniform(-0.45, 0.45, nant)
    subbands = [(256, 8e6, 1.65e9, 12, [0, 1, 3]), (1024, 32e6, 1.70e9, 5, None), (512, 16e6, 1.6e9, 9, [4, 2, 0, 1])]
    cfg = cfgmod.Config(nchan=256, lo=1.65e9, bandwidth=8e6, numffts=12, nsubint=nsub)
    fileoff = delays[:, 2] + spread / (2 * 8e6)     # fixed seconds for every subband
    for j in range(nant):
        cfg.antennas.append(cfgmod.AntennaEntry(f"A{j}", f"a{j}_sb%d.bin", tuple(delays[j]), float(fileoff[j])))
    streams = []
    for i, (nchan, bw, lo, nf, ants) in enumerate(subbands):
        cfg.subbands.append(cfgmod.SubbandEntry(nchan, bw, lo, nf, ants))
        F = 2 * nchan
        s = synth.make_data(nant, (nsub * nf + 1) * F, 40 + i)
        for j in range(nant):
            s[j].tofile(str(tmp_path / f"a{j}_sb{i}.bin"))
        streams.append(s)
    (tmp_path / "test.conf").write_text(cfgmod.format_config(cfg))
    outs = {}
    for mode in ([], ["--launch"]):
        r = subprocess.run([FXBATCH, "./test.conf", "--binary", *mode], cwd=tmp_path, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        outs[tuple(mode)] = [visio.read_binary(str(tmp_path / f"vis_sb{i}.bin")) for i in range(len(subbands))]
    for a, b in zip(outs[()], outs[("--launch",)]):
        assert np.array_equal(a["cross"], b["cross"]) and np.array_equal(a["cross_counts"], b["cross_counts"])
    for i, (nchan, bw, lo, nf, ants) in enumerate(subbands):
        idx = list(range(nant)) if ants is None else ants
        F = 2 * nchan
        vis = cnt = 0
        for s in range(nsub):
            p = orc.plan(nchan, lo, bw, origin=s * nf)
            buf = np.ascontiguousarray(streams[i][idx, s * nf * F // 2: (s * nf + nf + 1) * F // 2])
            r = orc.process_f32(p, buf, delays[idx], fileoff[idx], (s * nf + np.arange(nf)).astype(np.int32))
            vis = vis + r["vis"].astype(np.complex128)
            cnt = cnt + r["count"]
        d = outs[()][i]
        assert d["nant"] == len(idx) and d["nchan"] == nchan and d["flags"] == 1
        assert np.array_equal(d["cross_counts"], cnt)
        live = cnt > 0
        want = vis[live] / cnt[live, None, None]
        err = np.linalg.norm(d["cross"][live] - want, axis=-1) / np.linalg.norm(want, axis=-1)
        assert err.max() <= 1e-5
    # the last run was the --launch one
    log = (tmp_path / "runtime.fxbatch.log").read_text()
    assert "plain launches" in log and log.count("subint ") == nsub
