"""Output stage through the C ABI: device normalisation (processNormalize,
kernel.c:348-359) and the visibility files.

Bars: normalised visibilities equal the raw ones divided by their counts in
FP32 bit for bit (zero counts untouched); the binary file holds exactly the
normalised export; the text file holds the cross visibilities to its printed
6 decimals in the vis.out layout.
"""
import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, synth, visio

pytestmark = pytest.mark.gpu

LO = 1.65e9


def _engine_with_vis(nchan, nant, nf, seed, spread=2.4):
    p = orc.plan(nchan, LO, 32e6)
    rng = np.random.default_rng(seed)
    delays = np.stack([np.zeros(nant), rng.uniform(-1e-11, 1e-11, nant), rng.uniform(-1e-6, 1e-6, nant)], axis=1)
    # one antenna far outside the window: its baselines keep zero counts
    perturb = rng.uniform(-1.0, 1.0, nant)
    perturb[-1] = 1.0e6                      # every window of that antenna lies outside its buffer
    fileoff = delays[:, 2] + perturb * p.Ts
    data = synth.make_data(nant, (nf + 1) * 2 * nchan, seed)
    e = FxEngine(nant, nchan, LO, 32e6, prec="f16", max_ffts=nf, max_bytes_per_ant=data.shape[1])
    e.set_model(delays, fileoff)
    e.upload(data)
    e.reset()
    e.process(np.arange(nf, dtype=np.int32))
    return e


def test_device_normalisation():
    with _engine_with_vis(512, 6, 12, 91) as e:
        cross, cc, autos, ac = e.read_vis()
        ncross, ncc, nautos, nac = e.read_vis(normalize=True)
    assert np.array_equal(cc, ncc) and np.array_equal(ac, nac)
    assert (cc == 0).any() and (cc > 0).any()
    live = cc > 0
    want = cross.copy()
    want[live] = (cross[live].real / cc[live, None, None].astype(np.float32)) + \
        1j * (cross[live].imag / cc[live, None, None].astype(np.float32))
    assert np.array_equal(ncross, want.astype(np.complex64))
    alive = ac > 0
    wa = autos.copy()
    wa[alive] = (autos[alive].real / ac[alive, None, None].astype(np.float32)) + \
        1j * (autos[alive].imag / ac[alive, None, None].astype(np.float32))
    assert np.array_equal(nautos, wa.astype(np.complex64))


def test_binary_and_text_files(tmp_path):
    nchan, nant = 256, 4
    with _engine_with_vis(nchan, nant, 8, 92) as e:
        ncross, ncc, nautos, nac = e.read_vis(normalize=True)
        e.save_vis(str(tmp_path / "v.bin"), "binary", normalize=True)
        e.save_vis(str(tmp_path / "vis.out"), "text", normalize=True)
    d = visio.read_binary(str(tmp_path / "v.bin"))
    assert (d["nant"], d["nchan"], d["nbl"], d["flags"]) == (nant, nchan, 6, 1)
    assert d["bw"] == 32e6 and d["lo"] == LO
    assert np.array_equal(d["cross"], ncross) and np.array_equal(d["autos"], nautos)
    assert np.array_equal(d["cross_counts"], ncc) and np.array_equal(d["auto_counts"], nac)
    cols, vis, amp, ph = visio.read_text(str(tmp_path / "vis.out"), 6, nchan)
    assert np.array_equal(cols[:, 0], np.arange(nchan))
    assert np.allclose(cols[:, 1], (np.arange(nchan) + 0.5) / nchan * 32.0, atol=5e-7)
    # each printed component within half a unit of the 6th decimal
    assert np.abs(vis.real - ncross.real).max() <= 5.01e-7
    assert np.abs(vis.imag - ncross.imag).max() <= 5.01e-7
    # amplitude: float32(sqrt(double)) printed to 6 decimals
    ref = np.abs(ncross.astype(np.complex128))
    assert (np.abs(amp - ref) <= 2e-7 * ref + 1e-6).all()
