"""The correlator workloads the engine is measured and tested on.

Each workload fixes antennas, channels, sample rate, subintegration and the
seeded delay model (a quadratic polynomial per antenna in FFT-index units,
loadConfig.h:30).  LO = 1.65 GHz as in the reference configs (test.conf:4);
bandwidth = sample rate / 2 (real sampling).  The FFTs of a subint are
floor(subint * fs / F); a buffer holds one more FFT of samples, as
allocDataHost.c:21 sizes it.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from . import synth

LO = 1.65e9


@dataclass(frozen=True)
class Workload:
    name: str
    nant: int
    nchan: int
    fs: float                    # samples per second (real)
    subint_s: float = 0.01       # seconds per subintegration
    integration_s: float = 0.01  # seconds per integration
    delay_range: float = 2e-6    # |c2| bound (s)
    rate_max: float = 0.0        # |delay rate| bound (s/s)
    fringe_hz: float = 0.0       # |fringe rate| bound (Hz); adds rate = f / LO
    frac_delay: bool = False     # add a fractional-sample delay U[0,1) Ts
    prec: str = "f32"
    lo: float = LO

    @property
    def bw(self) -> float:
        return self.fs / 2.0

    @property
    def F(self) -> int:
        return 2 * self.nchan

    @property
    def Ts(self) -> float:
        return 1.0 / self.fs

    @property
    def nffts(self) -> int:
        """FFTs per subintegration."""
        return int(self.subint_s * self.fs) // self.F

    @property
    def nsubints(self) -> int:
        return max(1, int(round(self.integration_s / self.subint_s)))

    @property
    def subint_bytes(self) -> int:
        """allocDataHost.c:21 for 2-bit real dual-pol data."""
        return self.nchan * 2 * (self.nffts + 1) * 2 // 8 * 2

    @property
    def samples_per_subint(self) -> int:
        """Input samples (antennas x pols x samples) processed per subint."""
        return self.nant * 2 * self.nffts * self.F

    @property
    def nbl(self) -> int:
        return self.nant * (self.nant - 1) // 2

    @property
    def nbl_with_autos(self) -> int:
        return self.nant * (self.nant + 1) // 2

    def with_(self, **kw) -> "Workload":
        return replace(self, **kw)


def delay_model(w: Workload, seed: int):
    """(delays [nant,3], fileoffsets [nant]) for one subint starting at FFT 0."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 7]))
    tfft = w.F / w.fs
    c2 = rng.uniform(-w.delay_range, w.delay_range, w.nant)
    rate = rng.uniform(-w.rate_max, w.rate_max, w.nant)
    if w.fringe_hz:
        rate = rate + rng.uniform(-w.fringe_hz, w.fringe_hz, w.nant) / w.lo
    accel = rng.uniform(-1e-12, 1e-12, w.nant)          # s/s^2, tiny curvature
    c1 = rate * tfft
    c0 = accel * tfft * tfft
    fileoff = np.round(c2 / w.Ts) * w.Ts                 # net delay ~0 at FFT 0
    if w.frac_delay:
        c2 = c2 + rng.uniform(0.0, 1.0, w.nant) * w.Ts
    return np.stack([c0, c1, c2], axis=1), fileoff


def net_delay_samples(w: Workload, delays: np.ndarray, fileoff: np.ndarray, g: float) -> np.ndarray:
    tau = delays[:, 0] * g * g + delays[:, 1] * g + delays[:, 2]
    return (tau - fileoff) / w.Ts


def make_subint(w: Workload, seed: int, nffts: int | None = None):
    """(data uint8 [nant][bytes], delays, fileoffsets, iteration) of one subint."""
    nf = w.nffts if nffts is None else nffts
    delays, fileoff = delay_model(w, seed)
    d = np.round(net_delay_samples(w, delays, fileoff, nf / 2.0)).astype(np.int64)
    nsamp = (nf + 1) * w.F
    data = synth.make_data(w.nant, nsamp, seed, d)
    return data, delays, fileoff, np.arange(nf, dtype=np.int32)


def make_stream_subints(w: Workload, seed: int, nsub: int):
    """(data uint8 [nant][bytes], delays, fileoffsets) of nsub consecutive
    subints in one buffer (nsub * nffts FFTs plus one more FFT of samples):
    subint r is FFTs r*nffts ... r*nffts + nffts - 1, global indices from 0,
    each with its own correlated data (the common signal at the model's
    integer net delay in the middle of that subint)."""
    nf = w.nffts
    delays, fileoff = delay_model(w, seed)
    shifts = np.stack([np.round(net_delay_samples(w, delays, fileoff, r * nf + nf / 2.0)).astype(np.int64)
                       for r in range(nsub + 1)])
    data = synth.make_stream(w.nant, nsub + 1, nf * w.F, seed, shifts)
    return np.ascontiguousarray(data[:, : (nsub * nf + 1) * w.F // 2]), delays, fileoff


# The benchmark workloads: plain FP32 at 4 antennas, FP32 at 16 antennas with
# fringe rates and fractional delays, FP16 at 64 antennas, FP8 at 256 antennas.
C1 = Workload("c1-4ant-1024ch", nant=4, nchan=1024, fs=128e6, rate_max=1e-9)
C2 = Workload("c2-16ant-2048ch", nant=16, nchan=2048, fs=128e6, integration_s=0.1, fringe_hz=50.0,
              frac_delay=True)
C3 = Workload("c3-64ant-4096ch", nant=64, nchan=4096, fs=128e6, integration_s=1.0, rate_max=3e-6, prec="f16")
C5 = Workload("c5-256ant-1024ch", nant=256, nchan=1024, fs=64e6, integration_s=1.0, rate_max=1e-7, prec="f8")
# Eight uneven subbands (channels, sample rate, antennas) of a 16-antenna array,
# each correlating its own subset of the antennas.
C4_SUBBANDS = [(256, 16e6, 8), (512, 32e6, 10), (1024, 64e6, 12), (2048, 128e6, 16), (4096, 256e6, 16),
               (8192, 256e6, 14), (16384, 256e6, 16), (1024, 128e6, 9)]
C4 = [Workload(f"c4-sb{i}-{n}ch", nant=a, nchan=n, fs=fs) for i, (n, fs, a) in enumerate(C4_SUBBANDS)]

WORKLOADS = {w.name: w for w in (C1, C2, C3, C5)}
