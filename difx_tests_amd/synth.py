"""Seeded synthetic 2-bit dual-polarisation data in the reference's layout.

Signal model, per polarisation p: a common white Gaussian s_p and independent
unit noise n_{j,p} per antenna, v_{j,p}[q] = s_p[q + d_j] / 3 + n_{j,p}[q]
(correlation coefficient 0.1, sigma = sqrt(10/9)), with d_j the antenna's
integer net delay in samples at the middle of the subint.  Samples are
quantised at +-0.98 sigma to codes 0..3 (levels -3.3359, -1, +1, +3.3359 in
kernel.c:191-192) and packed 4 per byte: sample q in byte q/2, time slot
t = q % 2, polarisation p at bits 2*(2t+p) (kernel.c:195-200).

Every draw comes from numpy's PCG64 keyed by (seed, stream, antenna, pol), so
the same call yields the same bytes on any machine.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

THRESH = 0.98
SIGMA = float(np.sqrt(1.0 / 9.0 + 1.0))


def _rng(*key: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([int(k) & 0xFFFFFFFF for k in key]))


def quantise(v: np.ndarray, sigma: float = SIGMA) -> np.ndarray:
    """Gaussian samples -> 2-bit codes (0: < -t, 1: [-t, 0), 2: [0, t), 3: >= t)."""
    t = THRESH * sigma
    return (np.searchsorted(np.array([-t, 0.0, t], dtype=v.dtype), v, side="right")).astype(np.uint8)


def pack_codes(c0: np.ndarray, c1: np.ndarray) -> np.ndarray:
    """Codes of pol 0 and pol 1 (equal even length) -> packed bytes."""
    assert c0.shape == c1.shape and c0.size % 2 == 0
    return (c0[0::2] | (c1[0::2] << 2) | (c0[1::2] << 4) | (c1[1::2] << 6)).astype(np.uint8)


def unpack_codes(packed: np.ndarray, nsamp: int):
    """Inverse of pack_codes (brute-force bit extraction)."""
    b = packed.astype(np.uint8)
    q = np.arange(nsamp)
    byte = b[q >> 1]
    t = q & 1
    return (byte >> (4 * t)) & 3, (byte >> (4 * t + 2)) & 3


def make_data(nant: int, nsamp: int, seed: int, delays_samples=None, rho_common: float = 1.0 / 3.0,
              identical: bool = False) -> np.ndarray:
    """Packed [nant][nsamp/2] bytes; delays_samples[j] = integer shift d_j."""
    assert nsamp % 2 == 0
    d = np.zeros(nant, np.int64) if delays_samples is None else np.asarray(delays_samples, np.int64)
    margin = int(np.abs(d).max()) if nant else 0
    out = np.empty((nant, nsamp // 2), np.uint8)
    common = [_rng(seed, 1, 0, p).standard_normal(nsamp + 2 * margin, dtype=np.float32) for p in range(2)]
    for j in range(nant):
        codes = []
        for p in range(2):
            s = common[p][margin + d[j]: margin + d[j] + nsamp]
            if identical:
                v = s * np.float32(rho_common) + _rng(seed, 2, 0, p).standard_normal(nsamp, dtype=np.float32)
            else:
                v = s * np.float32(rho_common) + _rng(seed, 2, j, p).standard_normal(nsamp, dtype=np.float32)
            codes.append(quantise(v))
        out[j] = pack_codes(codes[0], codes[1])
    return out


def make_stream(nant: int, nchunks: int, chunk: int, seed: int, shifts, rho_common: float = 1.0 / 3.0,
                nthreads: int = 0) -> np.ndarray:
    """Packed [nant][nchunks*chunk/2] bytes of a long stream: the make_data
    model per chunk of `chunk` samples, chunk c with its own common signal,
    shifted for antenna j by shifts[c][j] samples (the integer net delay of
    the delay model at that chunk), and its own noise.  Chunks are generated
    in parallel threads (numpy's generators release the GIL), each keyed by
    (seed, chunk, antenna, pol), so the bytes do not depend on the threads."""
    assert chunk % 2 == 0
    sh = np.asarray(shifts, np.int64).reshape(nchunks, nant)
    out = np.empty((nant, nchunks * chunk // 2), np.uint8)
    t = np.float32(THRESH * SIGMA)
    rho = np.float32(rho_common)

    def work(c: int) -> None:
        margin = int(np.abs(sh[c]).max())
        common = [_rng(seed, 21, c, p).standard_normal(chunk + 2 * margin, dtype=np.float32) for p in range(2)]
        for j in range(nant):
            codes = []
            for p in range(2):
                s = common[p][margin + sh[c, j]: margin + sh[c, j] + chunk]
                v = s * rho + _rng(seed, 22, c * nant + j, p).standard_normal(chunk, dtype=np.float32)
                # quantise(): the number of thresholds -t, 0, t at or below v
                codes.append(((v >= -t).astype(np.uint8) + (v >= 0) + (v >= t)).astype(np.uint8))
            out[j, c * chunk // 2: (c + 1) * chunk // 2] = pack_codes(codes[0], codes[1])

    n = nthreads or min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max(1, n)) as ex:
        list(ex.map(work, range(nchunks)))
    return out


def random_bytes(nant: int, nbytes: int, seed: int) -> np.ndarray:
    """Uniform random packed bytes (every code equally likely)."""
    return _rng(seed, 3, 0, 0).integers(0, 256, size=(nant, nbytes), dtype=np.uint8)


# ---- complex 8-bit samples (COMPLEX 1, NBIT 8): one time sample is 4 bytes,
# pol 0 (re, im) then pol 1 (re, im) as int8 (gkernel.cu:279-287)

COMPLEX8_SCALE = 24.0        # int8 units per unit rms of each component


def make_data_complex8(nant: int, nsamp: int, seed: int, delays_samples=None, rho_common: float = 1.0 / 3.0,
                       identical: bool = False) -> np.ndarray:
    """[nant][4*nsamp] bytes (int8 stored as uint8) of complex Gaussian
    samples with the common-signal model of make_data, rounded and clipped
    to int8 at COMPLEX8_SCALE per unit rms."""
    d = np.zeros(nant, np.int64) if delays_samples is None else np.asarray(delays_samples, np.int64)
    margin = int(np.abs(d).max()) if nant else 0
    out = np.empty((nant, 4 * nsamp), np.uint8)
    common = [_rng(seed, 11, 0, p).standard_normal((nsamp + 2 * margin, 2), dtype=np.float32) for p in range(2)]
    scale = np.float32(COMPLEX8_SCALE / SIGMA)
    for j in range(nant):
        v = np.empty((nsamp, 4), np.float32)
        for p in range(2):
            s = common[p][margin + d[j]: margin + d[j] + nsamp]
            key = 0 if identical else j
            noise = _rng(seed, 12, key, p).standard_normal((nsamp, 2), dtype=np.float32)
            v[:, 2 * p: 2 * p + 2] = s * np.float32(rho_common) + noise
        q = np.clip(np.rint(v * scale), -128, 127).astype(np.int8)
        out[j] = q.reshape(-1).view(np.uint8)
    return out


def random_bytes_complex8(nant: int, nsamp: int, seed: int) -> np.ndarray:
    """Uniform random int8 components (as uint8), 4 bytes per time sample."""
    return _rng(seed, 13, 0, 0).integers(0, 256, size=(nant, 4 * nsamp), dtype=np.uint8)
