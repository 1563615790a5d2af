"""The DIFXFXV1 visibility file layout (reader against an independent writer)."""
import numpy as np
import pytest

from difx_tests_amd import visio


def test_binary_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    nant, nchan = 5, 16
    nbl = nant * (nant - 1) // 2
    cross = (rng.normal(size=(nbl, 4, nchan)) + 1j * rng.normal(size=(nbl, 4, nchan))).astype(np.complex64)
    autos = (rng.normal(size=(nant, 4, nchan)) + 1j * rng.normal(size=(nant, 4, nchan))).astype(np.complex64)
    cc = rng.integers(0, 9, nbl).astype(np.int32)
    ac = rng.integers(0, 9, nant).astype(np.int32)
    path = str(tmp_path / "v.bin")
    visio.write_binary(path, nant, nchan, 1, 32e6, 1.65e9, cc, ac, cross, autos)
    # layout by hand: 8 + 16 + 16 header bytes, then counts and data
    raw = open(path, "rb").read()
    assert raw[:8] == b"DIFXFXV1"
    assert np.array_equal(np.frombuffer(raw, "<i4", 4, 8), [nant, nchan, nbl, 1])
    assert np.array_equal(np.frombuffer(raw, "<f8", 2, 24), [32e6, 1.65e9])
    assert np.array_equal(np.frombuffer(raw, "<i4", nbl, 40), cc)
    off = 40 + 4 * (nbl + nant)
    assert np.array_equal(np.frombuffer(raw, "<c8", nbl * 4 * nchan, off).reshape(nbl, 4, nchan), cross)
    assert len(raw) == off + 8 * (nbl + nant) * 4 * nchan
    d = visio.read_binary(path)
    assert (d["nant"], d["nchan"], d["nbl"], d["flags"], d["bw"], d["lo"]) == (nant, nchan, nbl, 1, 32e6, 1.65e9)
    assert np.array_equal(d["cross"], cross) and np.array_equal(d["autos"], autos)
    assert np.array_equal(d["cross_counts"], cc) and np.array_equal(d["auto_counts"], ac)


def test_binary_rejects_other_files(tmp_path):
    p = tmp_path / "x.bin"
    p.write_bytes(b"NOTAVIS!" + bytes(64))
    with pytest.raises(ValueError):
        visio.read_binary(str(p))
