"""CPU checks of the drop-in boundary: the shared library loads without a GPU
and exports every function the public headers declare; the config grammar
matches the reference's parser (main.c:14-129)."""
import os
import re

import pytest

from difx_tests_amd import config, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return", "sizeof")))


@pytest.mark.parametrize("header", ["kernel.h", "difxfx.h"])
def test_library_exports_every_declared_symbol(header):
    lib = engine.load_library()
    declared = _declared_functions(header)
    assert len(declared) >= (28 if header == "kernel.h" else 20)
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_symbol_lists_match_headers():
    assert sorted(engine.KERNELH_SYMBOLS) == _declared_functions("kernel.h")
    assert sorted(engine.NATIVE_SYMBOLS) == _declared_functions("difxfx.h")


def test_reference_config_fixture_parses():
    # tests/golden/test4.conf is the reference's 4-antenna config (test4.conf)
    cfg = config.parse_config(os.path.join(GOLDEN, "test4.conf"), warn=None)
    assert (cfg.nbit, cfg.iscomplex, cfg.nchan, cfg.numffts, cfg.npol, cfg.nant) == (2, 0, 256, 3840, 2, 4)
    assert cfg.lo == 1650000000.0 and cfg.bandwidth == 32000000.0
    assert cfg.antennas[0].name == "Aa" and cfg.antennas[0].file == "aa.bin"
    assert cfg.antennas[3].delay == (3.9e-19, 9.1e-15, 129.9e-9) and cfg.antennas[3].fileoffset == 130e-9
    # allocDataHost.c:21: nchan * 2 * (numffts + 1) * 2 / 8 * 2
    assert cfg.subintbytes == 256 * 2 * 3841 * 2 // 8 * 2


def test_config_grammar_rules():
    txt = "nbit 2\nNchan 64\nlo 1e9\nbandwidth 8e6\nNUMFFTS 4\nFOO 3\nNANT 1\nA a.bin 0 0 1e-7 1e-7\n"
    cfg = config.parse_config_text(txt, warn=None)
    assert (cfg.nchan, cfg.numffts, cfg.nant, cfg.npol, cfg.iscomplex) == (64, 4, 1, 2, 0)
    with pytest.raises(config.ConfigError):          # empty line exits (main.c:37-40)
        config.parse_config_text("NBIT 2\n\nNCHAN 4\n", warn=None)
    with pytest.raises(config.ConfigError):          # missing value
        config.parse_config_text("NCHAN\n", warn=None)
    with pytest.raises(config.ConfigError):          # short antenna line
        config.parse_config_text("NANT 1\nA a.bin 0 0 1e-7\n", warn=None)
    again = config.parse_config_text(config.format_config(cfg), warn=None)
    assert again == cfg


def test_engine_fails_loudly_without_library(tmp_path):
    engine._lib_backup = engine._lib
    try:
        engine._lib = None
        with pytest.raises(engine.FxError):
            engine.load_library(str(tmp_path / "missing.so"))
    finally:
        engine._lib = engine._lib_backup


def test_subband_grammar_round_trip():
    from difx_tests_amd import config as cfgmod
    text = ("NBIT 2\nLO 1650000000\nNANT 3\nA a%d.bin 0 1e-12 1e-6 1e-6\nB b%d.bin 0 0 0 0\nC c%d.bin 0 0 2e-7 2e-7\n"
            "SUBBAND 256 8000000 1650000000 40 0,2\nSUBBAND 1024 32000000 1700000000 10 ALL\nNSUBINT 3\nPREC f16\n")
    cfg = cfgmod.parse_config_text(text, warn=None)
    assert cfg.nant == 3 and cfg.nsubint == 3 and cfg.prec == "F16"
    assert [(s.nchan, s.bandwidth, s.lo, s.numffts, s.antennas) for s in cfg.subbands] == \
        [(256, 8e6, 1.65e9, 40, [0, 2]), (1024, 32e6, 1.7e9, 10, None)]
    again = cfgmod.parse_config_text(cfgmod.format_config(cfg), warn=None)
    assert again == cfg


@pytest.mark.parametrize("kw,code", [
    (dict(nchan=1000), -1),            # not a power of two
    (dict(nchan=128), -4),             # FFT length 256 below the supported range
    (dict(nant=0), -1),
    (dict(bw=0.0), -1),
    (dict(lo=-1.0), -1),
    (dict(lo=2.2e9), -1),              # the reference's (int)lo needs lo < 2^31
    (dict(prec=7), -1),
    (dict(max_ffts=0), -1),
    (dict(iscomplex=2), -1),
])
def test_create_rejects_bad_parameters_before_any_device_call(kw, code):
    # fx_create checks its parameters before it looks for a GPU, so these
    # fail the same way with or without one
    import ctypes as C
    lib = engine.load_library()
    prm = dict(nant=4, nchan=1024, lo=1.65e9, bw=32e6, prec=0, max_ffts=16, max_bytes_per_ant=0, iscomplex=0)
    prm.update(kw)
    p = engine._Params(**prm)
    ctx = C.c_void_p(1)
    lib.fx_create.argtypes = [C.c_void_p, C.c_void_p]
    assert lib.fx_create(C.byref(p), C.byref(ctx)) == code
    assert not ctx.value
    lib.fx_last_error.restype = C.c_char_p
    assert lib.fx_last_error()


def test_kernelh_zero_ffts_leaves_accumulators():
    # numffts = 0: the reference's FFT loop (kernel.c:725) runs no iteration,
    # so nothing is accumulated; the pointer arrays are still copied out
    import ctypes as C
    import numpy as np
    lib = engine.load_library()
    nant, nchan = 3, 512
    nbl = nant * (nant - 1) // 2
    vis = np.arange(nbl * 4 * nchan * 2, dtype=np.float32).reshape(nbl, 4, nchan * 2)
    before = vis.copy()
    P = C.c_void_p
    rows = [(P * 4)(*[vis[b, q].ctypes.data for q in range(4)]) for b in range(nbl)]
    visp = (P * nbl)(*[C.cast(r, P) for r in rows])
    out = (P * nbl)()
    cnt = (C.c_int * nbl)(5, 6, 7)
    cnt_out = (C.c_int * nbl)()
    i = lambda v: C.byref(C.c_int(v))
    d = lambda v: C.byref(C.c_double(v))
    null = (P * nant)()
    lib.processAntennasAndBaselinePara.argtypes = [P] * 22 + [C.c_int]
    lib.processAntennasAndBaselinePara.restype = None
    lib.processAntennasAndBaselinePara(
        i(nant), i(0), i(2 * nchan), (C.c_double * nant)(), d(1 / 64e6), null, null, null, null, null,
        i(2), i(32), i(16), i(0), d(1.65e9), i(nchan), visp, cnt, (C.c_int * 1)(), d(32e6), out, cnt_out, 1)
    assert np.array_equal(vis, before)
    assert list(cnt) == [5, 6, 7] and list(cnt_out) == [5, 6, 7]
    assert [out[b] for b in range(nbl)] == [visp[b] for b in range(nbl)]
