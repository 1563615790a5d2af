"""Pins the CPU oracle (oracle/fx_oracle.c) against independent checks.

The reference ships no test vectors for this path (its only known answer,
main.c:443-444, needs the missing aa.bin), so every function of the oracle
is pinned here by something it does not share code with: brute-force bit
extraction, closed forms of the delay model and rotators, numpy's FFT and a
brute-force DFT, an einsum X-engine, and invariants.  Several tests also
plant a plausible mistake (on the CPU) and check that the comparison they
rely on would catch it.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import synth

LEVELS = np.array([-3.3359, -1.0, 1.0, 3.3359], np.float32)


# ---------------------------------------------------------------- unpack / LUT

def test_levels_known_answer():
    # kernel.c:191-192: {-HiMag, -1, 1, HiMag} with HiMag = (float)3.3359
    got = np.array([orc.lib().or_level(c) for c in range(4)], np.float32)
    assert np.array_equal(got, LEVELS)


def test_unpack_known_byte():
    # byte 0b11100100: pol0 t0 = 0, pol1 t0 = 1, pol0 t1 = 2, pol1 t1 = 3 (kernel.c:195-200)
    data = np.array([0b11100100, 0], np.uint8)
    a, b = orc.unpack(data, 0, 2)
    assert list(a) == [LEVELS[0], LEVELS[2]] and list(b) == [LEVELS[1], LEVELS[3]]
    a, b = orc.unpack(data, 1, 2)    # odd offset starts at the byte's second slot
    assert list(a) == [LEVELS[2], LEVELS[0]] and list(b) == [LEVELS[3], LEVELS[0]]


@pytest.mark.parametrize("offset", [0, 1, 2, 7, 1000, 1001])
def test_unpack_bruteforce(offset):
    rng = np.random.default_rng(offset)
    data = rng.integers(0, 256, 4096, dtype=np.uint8)
    F = 512
    a, b = orc.unpack(data, offset, F)
    c0, c1 = synth.unpack_codes(data, offset + F)
    assert np.array_equal(a, LEVELS[c0[offset:]]) and np.array_equal(b, LEVELS[c1[offset:]])


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(5)
    c0 = rng.integers(0, 4, 1000).astype(np.uint8)
    c1 = rng.integers(0, 4, 1000).astype(np.uint8)
    u0, u1 = synth.unpack_codes(synth.pack_codes(c0, c1), 1000)
    assert np.array_equal(u0, c0) and np.array_equal(u1, c1)


# ---------------------------------------------------------------- delay model

def test_station_delay_constant_and_linear():
    m, a, b = orc.station_delay([0.0, 0.0, 1.25e-7], 17)
    assert (m, a, b) == (1.25e-7, 0.0, 1.25e-7)
    c1, c2 = 3.0e-13, -4.0e-7
    for g in (0, 1, 155, 15624):
        m, a, b = orc.station_delay([0.0, c1, c2], g)
        assert a == pytest.approx(c1, rel=1e-9, abs=1e-24)
        assert b == pytest.approx(c1 * g + c2, rel=1e-12)
        assert m == pytest.approx(c1 * (g + 0.5) + c2, rel=1e-12)


def test_station_delay_quadratic_closed_form():
    # kernel.c:378-385 with tau(t) = c0 t^2 + c1 t + c2:
    #   a = c0 (2g+1) + c1,  b = tau(g) - c0/12,  mean = c0 (g^2 + g + 5/12) + c1 (g + 1/2) + c2
    c0, c1, c2 = 2.5e-16, -3.0e-12, 1.0e-6
    for g in (0, 3, 100, 2000):
        m, a, b = orc.station_delay([c0, c1, c2], g)
        assert a == pytest.approx(c0 * (2 * g + 1) + c1, rel=1e-9)
        assert b == pytest.approx(c0 * g * g + c1 * g + c2 - c0 / 12, rel=1e-12)
        assert m == pytest.approx(c0 * (g * g + g + 5 / 12) + c1 * (g + 0.5) + c2, rel=1e-12)
        # a planted mistake (true average c0/3 instead of 5 c0/12) is visible
        assert abs(m - (c0 * (g * g + g + 1 / 3) + c1 * (g + 0.5) + c2)) > 1e-3 * c0


@pytest.mark.parametrize("x,n", [(0.3, 0), (0.7, 1), (-0.3, 0), (-0.7, 0), (-1.2, 0), (-1.6, -1), (2.49, 2),
                                  (2.5, 3), (-2.5, -2)])
def test_offset_truncating_round(x, n):
    # kernel.c:737 (int)(x + 0.5) truncates toward zero (Q3); Ts = 0.25 keeps
    # every quantity exact in binary.
    Ts, F, g = 0.25, 8, 5
    v, o, fd = orc.offset(x * Ts, 0.0, Ts, g, F, 1000)
    assert v == 1 and o == g * F - n and fd == -(x - n) * Ts


def test_offset_edge_clamps_and_validity():
    Ts, F = 0.25, 8
    maxo = 3 * F
    # o == -1 -> 0 with fd + Ts (kernel.c:742-746)
    v, o, fd = orc.offset(1.0 * Ts, 0.0, Ts, 0, F, maxo)
    assert (v, o, fd) == (1, 0, 0.0 + Ts)
    # o == maxo + 1 -> maxo with fd - Ts (kernel.c:747-751); x = -1.5 truncates to n = -1
    v, o, fd = orc.offset(-1.5 * Ts, 0.0, Ts, 3, F, maxo)
    assert (v, o, fd) == (1, maxo, 0.5 * Ts - Ts)
    # x = -1.0 truncates to n = 0, so no clamp happens
    assert orc.offset(-1.0 * Ts, 0.0, Ts, 3, F, maxo) == (1, maxo, Ts)
    # o == -2 and o == maxo + 2 are invalid (kernel.c:752-755)
    assert orc.offset(2.0 * Ts, 0.0, Ts, 0, F, maxo)[0] == 0
    assert orc.offset(-2.5 * Ts, 0.0, Ts, 3, F, maxo)[0] == 0
    # the file offset is subtracted before dividing by Ts
    v, o, fd = orc.offset(0.75, 0.5, Ts, 2, F, maxo)
    assert (v, o) == (1, 2 * F - 1)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("F", [8, 64, 512, 2048, 8192])
def test_fft_matches_numpy(F):
    rng = np.random.default_rng(F)
    x = rng.standard_normal(F) + 1j * rng.standard_normal(F)
    ref = np.fft.fft(x)
    y64 = orc.fft(x, f64=True)
    assert np.linalg.norm(y64 - ref) / np.linalg.norm(ref) < 1e-13
    y32 = orc.fft(x.astype(np.complex64))
    assert np.linalg.norm(y32 - ref) / np.linalg.norm(ref) < 2e-6


@pytest.mark.parametrize("F", [16, 64])
def test_fft_matches_bruteforce_dft(F):
    rng = np.random.default_rng(F + 1)
    x = rng.standard_normal(F) + 1j * rng.standard_normal(F)
    n = np.arange(F)
    W = np.exp(-2j * np.pi * np.outer(n, n) / F)
    assert np.allclose(orc.fft(x, f64=True), W @ x, rtol=0, atol=1e-11)


def test_fft_single_tone_and_parseval():
    F, k0 = 2048, 37
    x = np.exp(2j * np.pi * k0 * np.arange(F) / F)
    y = orc.fft(x, f64=True)
    assert abs(y[k0] - F) < 1e-9 and np.max(np.abs(np.delete(y, k0))) < 1e-9
    rng = np.random.default_rng(2)
    z = (rng.standard_normal(F) + 1j * rng.standard_normal(F)).astype(np.complex64)
    Z = orc.fft(z)
    assert np.sum(np.abs(Z) ** 2) == pytest.approx(F * np.sum(np.abs(z.astype(np.complex128)) ** 2), rel=1e-5)


# ---------------------------------------------------------------- rotators

def _fringe_phase_turns(p, a, b):
    F, sub, stride, lo = p.F, p.sub, p.stride, p.lo
    n = np.arange(F)
    return ((n % sub) / F * a + (b - int(b))) * lo


def _fringe_closed_form(p, a, b):
    """exp(-2 pi i phase), phase = ((s + k*stride)/F * a + frac(b)) * lo [+ frac(lo)*int(b)].

    Sample n = k*sub + s; the coarse step uses k*stride (the reference's
    stepxoff, kernel.c:145) although blocks are sub = 2*stride samples long.
    """
    F, sub, stride, lo = p.F, p.sub, p.stride, p.lo
    n = np.arange(F)
    k, s = n // sub, n % sub
    ib = int(b)
    ph = ((s / F) * a + (b - ib)) * lo + ((k * stride) / F * a) * lo
    if p.fracLO:
        ph = ph + (lo - int(lo)) * ib
    return np.exp(-2j * np.pi * ph)


@pytest.mark.parametrize("nchan,a,b,lo", [(1024, 1.9e-10, 3.3e-7, 1.65e9), (256, -4.0e-11, -1.1e-6, 1.65e9),
                                          (2048, 7.0e-11, 2.0e-7, 1.65e9), (1024, 1.0e-10, 1.2, 1.65e9 + 0.5),
                                          (4096, 0.0, 5.0e-7, 1.65e9)])
def test_fringe_rotator_closed_form(nchan, a, b, lo):
    p = orc.plan(nchan, lo, 32e6)
    want = _fringe_closed_form(p, a, b)
    r64 = orc.fringe_rotator(p, a, b, f64=True)
    r32 = orc.fringe_rotator(p, a, b)
    # an FP64 phase of T turns is resolved to ~T * 2^-52 turns
    slack = 8e-16 * 2 * np.pi * np.max(np.abs(_fringe_phase_turns(p, a, b)))
    assert np.max(np.abs(r64 - want)) < 1e-8 + slack
    assert np.max(np.abs(r32 - want)) < 2e-6 + slack
    if a != 0.0:
        # planted mistake: coarse step of sub samples instead of stride
        n = np.arange(p.F)
        wrong = want * np.exp(-2j * np.pi * ((n // p.sub) * p.stride / p.F * a) * lo)
        assert np.max(np.abs(r32 - wrong)) > 1e-3


def test_fringe_rotator_constant_phase_when_a_is_zero():
    p = orc.plan(1024, 1.65e9, 64e6)
    r = orc.fringe_rotator(p, 0.0, 2.5e-7, f64=True)
    assert np.allclose(r, r[0], atol=1e-12)
    assert np.allclose(r[0], np.exp(-2j * np.pi * ((2.5e-7 * 1.65e9) % 1.0)), atol=1e-9)


@pytest.mark.parametrize("nchan,fd", [(1024, 3.1e-9), (256, -7.4e-9), (2048, 1.2e-8), (4096, -2.0e-9)])
def test_frac_rotator_closed_form(nchan, fd):
    bw = 64e6
    p = orc.plan(nchan, 1.65e9, bw)
    c = np.arange(nchan)
    want64 = np.exp(1j * 2 * np.pi * c * bw / nchan * fd)
    assert np.max(np.abs(orc.frac_rotator(p, fd, f64=True) - want64)) < 1e-10
    # the FP32 path narrows fd to f32 first (Q6)
    want32 = np.exp(1j * 2 * np.pi * c * bw / nchan * float(np.float32(fd)))
    assert np.max(np.abs(orc.frac_rotator(p, fd) - want32)) < 2e-6


def test_sincos_is_rounded_double():
    for x in (0.0, 1e-3, -0.5, 2.0, -6.2, 3.14159):
        s, c = orc.sincos(x)
        xf = float(np.float32(x))
        assert s == float(np.float32(np.sin(xf))) and c == float(np.float32(np.cos(xf)))


# ---------------------------------------------------------------- end to end

def _numpy_subint(p, data, delays, fileoff, iteration, split=1):
    """Independent numpy composition of the subint (FP64)."""
    nant = data.shape[0]
    F, nchan = p.F, p.nchan
    maxo = (len(iteration) * split - 1) * F
    chan = np.zeros((len(iteration), nant, 2, nchan), np.complex128)
    valid = np.zeros((len(iteration), nant), bool)
    c0all, c1all = zip(*[synth.unpack_codes(data[j], 2 * data.shape[1]) for j in range(nant)])
    for k, g in enumerate(iteration):
        for j in range(nant):
            c0, c1, c2 = delays[j]
            d = [c0 * t * t + c1 * t + c2 for t in (g, g + 0.5, g + 1.0)]
            a = d[2] - d[0]
            b = d[0] + (d[1] - (a * 0.5 + d[0])) / 3.0
            mean = a * 0.5 + b
            x = (mean - fileoff[j]) / p.Ts
            n = int(x + 0.5)
            fd = -(x - n) * p.Ts
            o = g * F - n
            if o == -1:
                o, fd = 0, fd + p.Ts
            if o == maxo + 1:
                o, fd = o - 1, fd - p.Ts
            if o < 0 or o > maxo:
                continue
            valid[k, j] = True
            rot = _fringe_closed_form(p, a, b)
            fr = np.exp(1j * 2 * np.pi * np.arange(nchan) * p.bw / nchan * fd)
            for pol, codes in enumerate((c0all[j], c1all[j])):
                u = LEVELS[codes[o:o + F]].astype(np.float64) * rot
                chan[k, j, pol] = np.fft.fft(u)[:nchan] * fr
    vis = []
    cnt = []
    for j in range(nant):
        for kk in range(j + 1, nant):
            m = valid[:, j] & valid[:, kk]
            v = np.einsum("fc,fc->c", chan[m, j, 0], np.conj(chan[m, kk, 0]))
            prods = [np.einsum("fc,fc->c", chan[m, j, l], np.conj(chan[m, kk, mm])) for l in (0, 1) for mm in (0, 1)]
            vis.append(prods)
            cnt.append(int(m.sum()))
            del v
    return np.array(vis), np.array(cnt), chan, valid


@pytest.mark.parametrize("nchan,nant,nf,seed", [(256, 4, 6, 1), (512, 3, 4, 2), (1024, 4, 3, 3)])
def test_process_matches_numpy_composition(nchan, nant, nf, seed):
    bw = 32e6
    p = orc.plan(nchan, 1.65e9, bw)
    F = 2 * nchan
    rng = np.random.default_rng(seed)
    data = synth.random_bytes(nant, (nf + 1) * F // 2, seed)
    delays = np.stack([rng.uniform(-1e-15, 1e-15, nant), rng.uniform(-3e-11, 3e-11, nant),
                       rng.uniform(-2e-7, 2e-7, nant)], axis=1)
    fileoff = np.round(delays[:, 2] / p.Ts) * p.Ts + rng.uniform(-1.4, 1.4, nant) * p.Ts
    it = np.arange(nf)
    r32 = orc.process_f32(p, data, delays, fileoff, it, want_chan=True)
    r64 = orc.process_f64(p, data, delays, fileoff, it, want_chan=True)
    vis, cnt, chan, valid = _numpy_subint(p, data, delays, fileoff, it)
    assert np.array_equal(r32["valid"].astype(bool), valid) and np.array_equal(r64["valid"].astype(bool), valid)
    assert np.array_equal(r32["count"], cnt)
    for r, tol in ((r64, 1e-9), (r32, 2e-5)):
        err = np.linalg.norm(r["chan"][valid] - chan[valid], axis=-1) / np.linalg.norm(chan[valid], axis=-1)
        assert err.max() < tol
        ve = np.linalg.norm(r["vis"] - vis, axis=-1) / np.maximum(np.linalg.norm(vis, axis=-1), 1e-30)
        assert ve.max() < tol


def test_identical_antennas_cross_equals_auto():
    bw, nchan, nf = 32e6, 256, 4
    p = orc.plan(nchan, 1.65e9, bw)
    one = synth.random_bytes(1, (nf + 1) * nchan, 9)
    data = np.repeat(one, 3, axis=0)
    delays = np.tile([0.0, 0.0, 1e-7], (3, 1))
    fileoff = np.full(3, 1e-7)
    r = orc.process_f64(p, data, delays, fileoff, np.arange(nf))
    for b in range(3):
        assert np.allclose(r["vis"][b], r["autos"][0], rtol=1e-12, atol=1e-6)
    # (p,p) autocorrelations are real and positive
    assert np.all(r["autos"][:, 0].real > 0) and np.max(np.abs(r["autos"][:, 0].imag)) < 1e-6 * np.max(r["autos"][:, 0].real)
    assert np.allclose(r["autos"][:, 1], np.conj(r["autos"][:, 2]))


def test_xstage_matches_einsum():
    rng = np.random.default_rng(4)
    nf, nant, nchan = 5, 5, 64
    chan = (rng.standard_normal((nf, nant, 2, nchan)) + 1j * rng.standard_normal((nf, nant, 2, nchan))).astype(np.complex64)
    valid = rng.integers(0, 2, (nf, nant)).astype(np.int32)
    vis, cnt = orc.xstage_f32(chan, valid)
    b = 0
    for j in range(nant):
        for k in range(j + 1, nant):
            m = (valid[:, j] & valid[:, k]).astype(bool)
            assert cnt[b] == m.sum()
            for l in (0, 1):
                for mm in (0, 1):
                    want = np.einsum("fc,fc->c", chan[m, j, l].astype(np.complex128), np.conj(chan[m, k, mm]))
                    assert np.allclose(vis[b, 2 * l + mm], want, rtol=1e-5, atol=1e-4)
            b += 1


def test_fstage_threads_match_single_thread():
    bw, nchan, nf, nant = 32e6, 512, 6, 3
    p = orc.plan(nchan, 1.65e9, bw)
    data = synth.random_bytes(nant, (nf + 1) * nchan, 3)
    delays = np.tile([0.0, 1e-11, 1e-7], (nant, 1))
    fileoff = np.full(nant, 1e-7)
    a, va = orc.fstage_f32(p, data, delays, fileoff, np.arange(nf), nthreads=1)
    b, vb = orc.fstage_f32(p, data, delays, fileoff, np.arange(nf), nthreads=4)
    assert np.array_equal(a, b) and np.array_equal(va, vb)


# ---- buffer origin (streaming subints; origin 0 is the reference)

def test_offset_origin_zero_is_reference():
    rng = np.random.default_rng(11)
    for _ in range(200):
        mean, fo = rng.uniform(-3e-6, 3e-6, 2)
        g = int(rng.integers(0, 50))
        a = orc.offset(mean, fo, 1 / 64e6, g, 2048, 49 * 2048)
        b = orc.offset(mean, fo, 1 / 64e6, g, 2048, 49 * 2048, origin=0)
        assert a == b


def test_offset_origin_shifts_window_and_edges():
    # the same net delay relative to a buffer starting at FFT 7: the offset is
    # 7*F smaller and the clamps act at the buffer's own edges
    Ts, F, nf = 1 / 64e6, 2048, 10
    maxo = (nf - 1) * F
    for g in range(7, 7 + nf):
        for x in (-0.4, 0.6, -1.0, 1.4):           # net delay in samples
            mean = x * Ts
            v0, o0, fd0 = orc.offset(mean, 0.0, Ts, g, F, (g + 1) * F * 4)
            v1, o1, fd1 = orc.offset(mean, 0.0, Ts, g, F, maxo, origin=7)
            n = int(x + 0.5)
            if g == 7 and o0 - 7 * F == -1:
                assert (v1, o1) == (1, 0) and fd1 == fd0 + Ts       # low-edge clamp of the buffer
            elif g == 7 + nf - 1 and o0 - 7 * F == maxo + 1:
                assert (v1, o1) == (1, maxo) and fd1 == fd0 - Ts    # high-edge clamp
            else:
                assert o1 == o0 - 7 * F and fd1 == fd0 and o1 == (g - 7) * F - n
                assert v1 == (0 <= o1 <= maxo)


def test_streaming_subints_equal_whole_stream_in_the_interior():
    # three subints cut from one stream (each buffer holds NF + 1 FFTs of
    # samples, allocDataHost.c:21) with global FFT indices and origins give the
    # whole-stream spectra bit for bit wherever both place the window alike
    nchan, nant, nf, S = 256, 3, 6, 3
    F = 2 * nchan
    p0 = orc.plan(nchan, 1.65e9, 16e6)
    rng = np.random.default_rng(12)
    delays = np.stack([np.zeros(nant), rng.uniform(-2e-10, 2e-10, nant), rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-0.45, 0.45, nant) * p0.Ts
    stream = synth.random_bytes(nant, (S * nf + 1) * F // 2, 12)
    whole = orc.process_f32(p0, stream, delays, fileoff, np.arange(S * nf, dtype=np.int32), want_chan=True)
    compared = 0
    for s in range(S):
        ps = orc.plan(nchan, 1.65e9, 16e6, origin=s * nf)
        buf = np.ascontiguousarray(stream[:, s * nf * F // 2: (s * nf + nf + 1) * F // 2])
        it = (s * nf + np.arange(nf)).astype(np.int32)
        part = orc.process_f32(ps, buf, delays, fileoff, it, want_chan=True)
        for k in range(nf):
            g = s * nf + k
            for j in range(nant):
                same = part["offset"][k, j] + s * nf * F == whole["offset"][g, j]
                if part["valid"][k, j] and whole["valid"][g, j] and same:
                    assert np.array_equal(part["chan"][k, j], whole["chan"][g, j])
                    compared += 1
    assert compared >= S * nf * nant - 2 * S * nant


# ---- complex 8-bit input

def test_complex_plan_geometry():
    p = orc.plan(1024, 1.65e9, 32e6, complex8=True)
    assert (p.nchan, p.F, p.stride, p.sub, p.nstep, p.nfstep, p.iscomplex) == (1024, 1024, 32, 64, 16, 32, 1)
    assert p.Ts == 1.0 / 32e6
    r = orc.plan(1024, 1.65e9, 32e6)
    assert (r.F, r.nstep, r.iscomplex, r.Ts) == (2048, 32, 0, 1.0 / 64e6)
    assert orc.subint_bytes_complex(1024, 10) == 1024 * 11 * 4


def test_unpack8complex_brute_force():
    rng = np.random.default_rng(21)
    raw = rng.integers(0, 256, 4 * 300, dtype=np.uint8)
    x0, x1 = orc.unpack8complex(raw, 17, 200)
    v = raw.view(np.int8).reshape(300, 4).astype(np.float32)
    assert np.array_equal(x0, (v[17:217, 0] + 1j * v[17:217, 1]).astype(np.complex64))
    assert np.array_equal(x1, (v[17:217, 2] + 1j * v[17:217, 3]).astype(np.complex64))


def test_complex_fstage_matches_numpy_composition():
    # independent composition: int8 pairs -> x * rotator -> numpy FFT -> every
    # bin times the fractional rotator
    nchan, nant, nf = 256, 2, 3
    p = orc.plan(nchan, 1.65e9, 16e6, complex8=True)
    rng = np.random.default_rng(22)
    delays = np.stack([np.zeros(nant), rng.uniform(-1e-10, 1e-10, nant), rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    fileoff = delays[:, 2] + rng.uniform(-0.4, 0.4, nant) * p.Ts
    data = synth.random_bytes_complex8(nant, (nf + 1) * nchan, 22)
    it = np.arange(nf, dtype=np.int32)
    ref = orc.process_f32(p, data, delays, fileoff, it, want_chan=True)
    ref64 = orc.process_f64(p, data, delays, fileoff, it, want_chan=True)
    for k in range(nf):
        for j in range(nant):
            m, a, b = orc.station_delay(delays[j], k)
            v, o, fd = orc.offset(m, fileoff[j], p.Ts, k, p.F, (nf - 1) * p.F)
            assert v == ref["valid"][k, j] and o == ref["offset"][k, j]
            x0, x1 = orc.unpack8complex(data[j], o, p.F)
            rot = orc.fringe_rotator(p, a, b, f64=True)
            fr = orc.frac_rotator(p, fd, f64=True)
            for pol, x in enumerate((x0, x1)):
                want = fr * np.fft.fft(x.astype(np.complex128) * rot)
                got = ref["chan"][k, j, pol]
                assert np.linalg.norm(got - want) / np.linalg.norm(want) < 2e-6
                assert np.linalg.norm(ref64["chan"][k, j, pol] - want) / np.linalg.norm(want) < 1e-12


def test_complex_identical_antennas_cross_equals_auto():
    nchan, nf = 128, 4
    p = orc.plan(nchan, 1.65e9, 16e6, complex8=True)
    delays = np.tile([0.0, 0.0, 3e-7], (2, 1))
    fileoff = np.full(2, 3e-7)
    data = synth.make_data_complex8(2, (nf + 1) * nchan, 23, identical=True)
    r = orc.process_f64(p, data, delays, fileoff, np.arange(nf, dtype=np.int32))
    assert r["count"][0] == nf
    assert np.allclose(r["vis"][0], r["autos"][0], rtol=1e-12, atol=1e-9)
