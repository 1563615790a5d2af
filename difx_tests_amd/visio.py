"""Readers of the visibility files the engine writes (fx_save_vis).

Binary "DIFXFXV1" (little endian): magic[8], int32 nant, nchan, nbl, flags,
double bw, lo, int32 cross_counts[nbl], int32 auto_counts[nant],
complex64 cross[nbl][4][nchan], complex64 autos[nant][4][nchan].
Text: the reference's vis.out (kernel.c:608-617).
"""
from __future__ import annotations

import numpy as np

MAGIC = b"DIFXFXV1"


def read_binary(path: str) -> dict:
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a DIFXFXV1 visibility file")
        nant, nchan, nbl, flags = np.fromfile(f, "<i4", 4).tolist()
        bw, lo = np.fromfile(f, "<f8", 2).tolist()
        cc = np.fromfile(f, "<i4", nbl)
        ac = np.fromfile(f, "<i4", nant)
        cross = np.fromfile(f, "<c8", nbl * 4 * nchan).reshape(nbl, 4, nchan)
        autos = np.fromfile(f, "<c8", nant * 4 * nchan).reshape(nant, 4, nchan)
        if f.read(1):
            raise ValueError(f"{path}: trailing bytes")
    return dict(nant=nant, nchan=nchan, nbl=nbl, flags=flags, bw=bw, lo=lo, cross_counts=cc, auto_counts=ac,
                cross=cross, autos=autos)


def write_binary(path: str, nant, nchan, flags, bw, lo, cross_counts, auto_counts, cross, autos) -> None:
    """The same layout from numpy arrays (used to pin the reader)."""
    nbl = nant * (nant - 1) // 2
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([nant, nchan, nbl, flags], "<i4").tofile(f)
        np.array([bw, lo], "<f8").tofile(f)
        np.asarray(cross_counts, "<i4").tofile(f)
        np.asarray(auto_counts, "<i4").tofile(f)
        np.asarray(cross, "<c8").reshape(nbl, 4, nchan).tofile(f)
        np.asarray(autos, "<c8").reshape(nant, 4, nchan).tofile(f)


def read_text(path: str, nbl: int, nchan: int):
    """(channel/frequency columns [nchan, 2], cross [nbl, 4, nchan] complex, amplitude, phase)."""
    a = np.loadtxt(path, ndmin=2)
    assert a.shape == (nchan, 2 + nbl * 16), a.shape
    v = a[:, 2:].reshape(nchan, nbl, 4, 4)
    vis = (v[..., 0] + 1j * v[..., 1]).transpose(1, 2, 0)
    return a[:, :2], vis, v[..., 2].transpose(1, 2, 0), v[..., 3].transpose(1, 2, 0)
