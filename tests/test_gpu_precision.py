"""Per-dtype error statements against the FP64 oracle at the benchmark shapes.

64 antennas x 4096 channels (FP16 workload) and 256 antennas x 1024 channels
(FP8 workload), correlated 2-bit data (common signal, rho = 0.1), the same
subint through the FP32, FP16 and FP8 X-engines.  Errors per class:

  cross-par   cross baselines, parallel hands (products 0 and 3)
  cross-x     cross baselines, cross hands (products 1 and 2)
  auto-par    autocorrelations, parallel hands
  auto-x      autocorrelations, cross hands

with two metrics: per (baseline, product) ||V - V64|| / ||V64|| over channels
(max and rms over the class), and per element |V - V64| / sqrt(A_i A_j) with
the FP64 parallel-hand autocorrelations of the two signals (max and rms).
The statement is printed as one JSON line per workload; FX_PRECISION_NFFTS
sets the FFTs per subint (default 16).

Bars: FP32 per (baseline, product) <= 1e-5; FP16 <= 1e-3 for every class
(the FP16 target); FP8 <= 0.1 for crosses and cross-hand autos, <= 0.02 for
parallel-hand autos (e4m3 keeps 3 mantissa bits: ~5% rms per product).
"""
import json
import os

import numpy as np
import pytest

import pyoracle as orc
from difx_tests_amd import FxEngine, workloads

pytestmark = pytest.mark.gpu

NFFTS = int(os.environ.get("FX_PRECISION_NFFTS", "16"))


def _classes(cross, autos, ref, nant):
    rv, ra = ref["vis"], ref["autos"]
    # FP64 parallel-hand autos per (antenna, pol, channel)
    apow = np.stack([ra[:, 0].real, ra[:, 3].real], axis=1)          # [ant, pol, chan]
    jj, kk = np.triu_indices(nant, 1)
    out = {}
    for name, V, R, idx in (("cross-par", cross, rv, [0, 3]), ("cross-x", cross, rv, [1, 2]),
                            ("auto-par", autos, ra, [0, 3]), ("auto-x", autos, ra, [1, 2])):
        v, r = V[:, idx], R[:, idx]
        rel = np.linalg.norm(v - r, axis=-1) / np.maximum(np.linalg.norm(r, axis=-1), 1e-300)
        if name.startswith("cross"):
            pi, pj = jj, kk
        else:
            pi = pj = np.arange(nant)
        l = np.array([p // 2 for p in idx])
        m = np.array([p % 2 for p in idx])
        norm = np.sqrt(apow[pi][:, l] * apow[pj][:, m])                 # [nbl, 2, chan]
        el = np.abs(v - r) / np.maximum(norm, 1e-300)
        out[name] = {"bp_max": float(rel.max()), "bp_rms": float(np.sqrt(np.mean(rel ** 2))),
                     "el_max": float(el.max()), "el_rms": float(np.sqrt(np.mean(el ** 2)))}
    return out


def _run(w, prec, data, delays, fileoff, it):
    with FxEngine(w.nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=len(it), max_bytes_per_ant=data.shape[1]) as e:
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        e.process(it)
        return e.read_vis()


@pytest.mark.parametrize("wname,precs", [("c3-64ant-4096ch", ("f32", "f16", "f8")),
                                         ("c5-256ant-1024ch", ("f32", "f16", "f8"))])
def test_error_statement(wname, precs):
    w = workloads.WORKLOADS[wname]
    data, delays, fileoff, it = workloads.make_subint(w, 20261203, nffts=NFFTS)
    p = orc.plan(w.nchan, w.lo, w.bw)
    ref = orc.process_f64(p, data, delays, fileoff, it)
    stmt = {"workload": w.name, "nffts": NFFTS, "reference": "FP64 oracle"}
    res = {}
    for prec in precs:
        cross, cc, autos, ac = _run(w, prec, data, delays, fileoff, it)
        assert np.array_equal(cc, ref["count"]) and np.array_equal(ac, ref["autocount"])
        res[prec] = (cross, autos)
        stmt[prec] = _classes(cross, autos, ref, w.nant)
    if "f32" in res and "f16" in res:
        # FP16 against the FP32 GPU path, same metric
        f32ref = {"vis": res["f32"][0].astype(np.complex128), "autos": res["f32"][1].astype(np.complex128)}
        stmt["f16_vs_f32_gpu"] = _classes(res["f16"][0], res["f16"][1], f32ref, w.nant)
    print("PRECISION " + json.dumps(stmt), flush=True)
    for c, v in stmt["f32"].items():
        assert v["bp_max"] <= 1e-5, (c, v)
    for c, v in stmt["f16"].items():
        assert v["bp_max"] <= 1e-3, (c, v)
    for c, v in stmt["f8"].items():
        assert v["bp_max"] <= (0.02 if c == "auto-par" else 0.1), (c, v)
