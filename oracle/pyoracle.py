"""ctypes bindings of the CPU oracle (oracle/fx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libfxoracle.so")


class OrPlan(C.Structure):
    _fields_ = [("nchan", C.c_int), ("F", C.c_int), ("stride", C.c_int), ("sub", C.c_int), ("nstep", C.c_int),
                ("nfstep", C.c_int), ("Ts", C.c_double), ("lo", C.c_double), ("bw", C.c_double),
                ("fracLO", C.c_int), ("iscomplex", C.c_int), ("origin", C.c_int)]


_lib = None


def build() -> str:
    os.makedirs(os.path.dirname(SO), exist_ok=True)
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    os.path.join(HERE, "fx_oracle.c"), "-o", SO, "-lm"], check=True)
    return SO


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = C.CDLL(SO)
        P = C.c_void_p
        I = C.c_int
        D = C.c_double
        L.or_plan_init.argtypes = [C.POINTER(OrPlan), I, D, D]
        L.or_plan_init_complex.argtypes = [C.POINTER(OrPlan), I, D, D]
        L.or_subint_bytes_complex.argtypes = [I, I]
        L.or_subint_bytes_complex.restype = C.c_longlong
        L.or_unpack8complex.argtypes = [P, I, I, P, P]
        L.or_station_delay.argtypes = [P, I, P, P, P]
        L.or_offset.argtypes = [D, D, D, I, I, I, P, P]
        L.or_offset_origin.argtypes = [D, D, D, I, I, I, I, P, P]
        L.or_unpack2bit.argtypes = [P, I, I, P, P]
        L.or_level.argtypes = [I]
        L.or_level.restype = C.c_float
        L.or_fringe_rotator_f32.argtypes = [C.POINTER(OrPlan), D, D, P]
        L.or_fringe_rotator_f64.argtypes = [C.POINTER(OrPlan), D, D, P]
        L.or_frac_rotator_f32.argtypes = [C.POINTER(OrPlan), D, P]
        L.or_frac_rotator_f64.argtypes = [C.POINTER(OrPlan), D, P]
        L.or_fft_f32.argtypes = [I, P, P]
        L.or_fft_f64.argtypes = [I, P, P]
        L.or_sincos_f32.argtypes = [C.c_float, P, P]
        L.or_process_f32.argtypes = [C.POINTER(OrPlan), I, I, I, P, P, P, P, P, P, P, P, P, P, P, P]
        L.or_process_f64.argtypes = [C.POINTER(OrPlan), I, I, I, P, P, P, P, P, P, P, P, P, P]
        L.or_fstage_f32.argtypes = [C.POINTER(OrPlan), I, I, I, P, P, P, P, P, P, I]
        L.or_xstage_f32.argtypes = [I, I, I, P, P, P, P, I]
        L.or_subint_bytes.argtypes = [I, I]
        L.or_subint_bytes.restype = C.c_longlong
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def plan(nchan: int, lo: float, bw: float, origin: int = 0, complex8: bool = False) -> OrPlan:
    """origin: global FFT index of the data buffer's first FFT (0: the reference);
    complex8: 8-bit complex input (F = nchan, Ts = 1/bw)."""
    p = OrPlan()
    rc = (lib().or_plan_init_complex if complex8 else lib().or_plan_init)(C.byref(p), nchan, lo, bw)
    if rc:
        raise ValueError(f"or_plan_init failed ({rc})")
    p.origin = origin
    return p


def _data_ptrs(data: np.ndarray):
    rows = [np.ascontiguousarray(data[j]) for j in range(data.shape[0])]
    arr = (C.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    return arr, rows


def process_f32(p: OrPlan, data, delays, fileoff, iteration, split=1, want_chan=False):
    """FP32 restatement of processAntennasAndBaselinePara.  Returns dict."""
    nant = data.shape[0]
    it = np.ascontiguousarray(iteration, np.int32)
    nf = len(it)
    nbl = nant * (nant - 1) // 2
    vis = np.zeros((max(nbl, 1), 4, p.nchan), np.complex64)
    cnt = np.zeros(max(nbl, 1), np.int32)
    autos = np.zeros((nant, 4, p.nchan), np.complex64)
    acnt = np.zeros(nant, np.int32)
    chan = np.zeros((nf, nant, 2, p.nchan), np.complex64) if want_chan else None
    valid = np.zeros((nf, nant), np.int32)
    offs = np.zeros((nf, nant), np.int32)
    fds = np.zeros((nf, nant), np.float64)
    ptrs, keep = _data_ptrs(data)
    d = np.ascontiguousarray(delays, np.float64)
    o = np.ascontiguousarray(fileoff, np.float64)
    rc = lib().or_process_f32(C.byref(p), nant, nf, split, _p(d), _p(o), _p(it), ptrs, _p(vis), _p(cnt),
                              _p(autos), _p(acnt), _p(chan), _p(valid), _p(offs), _p(fds))
    if rc:
        raise RuntimeError("or_process_f32 failed")
    return dict(vis=vis[:nbl], count=cnt[:nbl], autos=autos, autocount=acnt, chan=chan, valid=valid,
                offset=offs, fracdelay=fds)


def process_f64(p: OrPlan, data, delays, fileoff, iteration, split=1, want_chan=False):
    nant = data.shape[0]
    it = np.ascontiguousarray(iteration, np.int32)
    nf = len(it)
    nbl = nant * (nant - 1) // 2
    vis = np.zeros((max(nbl, 1), 4, p.nchan), np.complex128)
    cnt = np.zeros(max(nbl, 1), np.int32)
    autos = np.zeros((nant, 4, p.nchan), np.complex128)
    acnt = np.zeros(nant, np.int32)
    chan = np.zeros((nf, nant, 2, p.nchan), np.complex128) if want_chan else None
    valid = np.zeros((nf, nant), np.int32)
    ptrs, keep = _data_ptrs(data)
    d = np.ascontiguousarray(delays, np.float64)
    o = np.ascontiguousarray(fileoff, np.float64)
    rc = lib().or_process_f64(C.byref(p), nant, nf, split, _p(d), _p(o), _p(it), ptrs, _p(vis), _p(cnt),
                              _p(autos), _p(acnt), _p(chan), _p(valid))
    if rc:
        raise RuntimeError("or_process_f64 failed")
    return dict(vis=vis[:nbl], count=cnt[:nbl], autos=autos, autocount=acnt, chan=chan, valid=valid)


def fstage_f32(p: OrPlan, data, delays, fileoff, iteration, split=1, nthreads=1):
    nant = data.shape[0]
    it = np.ascontiguousarray(iteration, np.int32)
    nf = len(it)
    chan = np.zeros((nf, nant, 2, p.nchan), np.complex64)
    valid = np.zeros((nf, nant), np.int32)
    ptrs, keep = _data_ptrs(data)
    d = np.ascontiguousarray(delays, np.float64)
    o = np.ascontiguousarray(fileoff, np.float64)
    rc = lib().or_fstage_f32(C.byref(p), nant, nf, split, _p(d), _p(o), _p(it), ptrs, _p(chan), _p(valid),
                             nthreads)
    if rc:
        raise RuntimeError("or_fstage_f32 failed")
    return chan, valid


def xstage_f32(chan, valid, nthreads=1):
    nf, nant, _, nchan = chan.shape
    nbl = nant * (nant - 1) // 2
    vis = np.zeros((max(nbl, 1), 4, nchan), np.complex64)
    cnt = np.zeros(max(nbl, 1), np.int32)
    c = np.ascontiguousarray(chan, np.complex64)
    v = np.ascontiguousarray(valid, np.int32)
    lib().or_xstage_f32(nant, nchan, nf, _p(c), _p(v), _p(vis), _p(cnt), nthreads)
    return vis[:nbl], cnt[:nbl]


def station_delay(c, g):
    mean, a, b = C.c_double(), C.c_double(), C.c_double()
    cc = np.ascontiguousarray(c, np.float64)
    lib().or_station_delay(_p(cc), g, C.byref(mean), C.byref(a), C.byref(b))
    return mean.value, a.value, b.value


def offset(mean, fileoff, Ts, g, F, maxoffset, origin=None):
    o, fd = C.c_int(), C.c_double()
    if origin is None:
        v = lib().or_offset(mean, fileoff, Ts, g, F, maxoffset, C.byref(o), C.byref(fd))
    else:
        v = lib().or_offset_origin(mean, fileoff, Ts, g, origin, F, maxoffset, C.byref(o), C.byref(fd))
    return v, o.value, fd.value


def unpack8complex(data_row, offset_, F):
    """(pol 0, pol 1) complex64 samples [F] starting at sample offset_."""
    a = np.zeros(F, np.complex64)
    b = np.zeros(F, np.complex64)
    d = np.ascontiguousarray(data_row, np.uint8)
    lib().or_unpack8complex(_p(d), offset_, F, _p(a), _p(b))
    return a, b


def subint_bytes_complex(nchan, numffts):
    return int(lib().or_subint_bytes_complex(nchan, numffts))


def unpack(data_row, offset_, F):
    a = np.zeros(F, np.float32)
    b = np.zeros(F, np.float32)
    d = np.ascontiguousarray(data_row, np.uint8)
    lib().or_unpack2bit(_p(d), offset_, F, _p(a), _p(b))
    return a, b


def fringe_rotator(p: OrPlan, a, b, f64=False):
    out = np.zeros(p.F, np.complex128 if f64 else np.complex64)
    (lib().or_fringe_rotator_f64 if f64 else lib().or_fringe_rotator_f32)(C.byref(p), a, b, _p(out))
    return out


def frac_rotator(p: OrPlan, fd, f64=False):
    out = np.zeros(p.nchan, np.complex128 if f64 else np.complex64)
    (lib().or_frac_rotator_f64 if f64 else lib().or_frac_rotator_f32)(C.byref(p), fd, _p(out))
    return out


def fft(x, f64=False):
    x = np.ascontiguousarray(x, np.complex128 if f64 else np.complex64)
    out = np.zeros_like(x)
    (lib().or_fft_f64 if f64 else lib().or_fft_f32)(len(x), _p(x), _p(out))
    return out


def sincos(x: float):
    s, c = C.c_float(), C.c_float()
    lib().or_sincos_f32(x, C.byref(s), C.byref(c))
    return s.value, c.value
