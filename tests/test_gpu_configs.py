"""GPU parity on the benchmark workloads themselves, at their full sizes,
through the C ABI, against the FP64 CPU oracle (BPMIteration.cpp:53-177 and
ScreenOutput.cpp:184-302 restated in oracle/).

Tolerances (relative L2 per output (wavelength, wavevector, polarisation)):
  FP64 path: 1e-10      FP32 path: 1e-4
Workloads:
  twisted nematic 64x64x20           (tests/test_gpu_parity.py)
  random-smooth 512x512x100          FP64 one wavelength at a time and the FP32
                                     graph-replayed batch, both against the oracle
  droplet 256x256x80, 294 props      FP32 graph batch and FP64 batch (mask formula
                                     weights), 3 of the 21 wavelengths (all 7
                                     wavevectors, both polarisations) against the
                                     oracle run on just those wavelengths
  cholesteric 1024x1024x200          FP64 and FP32 against the oracle
  serving trace over the four mesh sizes 128^2x40 .. 1024^2x200: every job the
                                     same in the mix as alone (bit for bit)
The oracle runs on the host's cores (about 5 s for 512^2x100, 25 s for the 42
droplet propagations, 45 s for 1024^2x200 on 16 threads).
"""
import json
import os
import sys

import numpy as np
import pytest

import mask_weights
import oracle
from nemaktis_amd import bpm_backend, fields, serving

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-4


def _rel_all(a, b):
    """relative L2 error of every output (wavelength, wavevector, polarisation)"""
    a = a.reshape(a.shape[0], a.shape[1], 2, -1)
    b = b.reshape(b.shape[0], b.shape[1], 2, -1)
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def _say(*a):
    print("[gpu config]", *a, flush=True)


def _context_run(s, lc, mask=None, formula="", **kw):
    shape = fields.output_shape(s)
    with bpm_backend.Context(json.dumps(s), lc, int(np.prod(shape)), mask_vals=mask, mask_formula=formula or "",
                             **kw) as ctx:
        ctx.propagate()
        return ctx.download(np.empty(shape, np.complex128))


def test_random512_full_fp64_and_fp32_against_oracle():
    s, lc, _, _ = fields.workload("random512")
    _say("random512: oracle")
    want = oracle.propagate(s, lc)
    _say("random512: FP64 one-shot")
    e64 = bpm_backend.propagate(json.dumps(s), lc)
    err64 = _rel_all(e64, want)
    _say("random512: FP32 graph batch")
    e32 = _context_run(s, lc, precision=32, wl_batch=0, use_graph=True)
    err32 = _rel_all(e32, want)
    _say("random512: FP64 max %.3e, FP32 max %.3e" % (err64.max(), err32.max()))
    assert err64.max() < TOL64, err64
    assert err32.max() < TOL32, err32


def test_droplet256_sweep_against_oracle_on_three_wavelengths():
    s, lc, mask, formula = fields.workload("droplet256")
    ic = s["Physics settings"]["Initial conditions"]
    dims, sp = ic["Mesh dimensions"], ic["Mesh spacings"]
    wl = s["Physics settings"]["Coefficients"]["Wavelengths"]
    pick = [0, 10, 20]
    sub = json.loads(json.dumps(s))
    sub["Physics settings"]["Coefficients"]["Wavelengths"] = [wl[i] for i in pick]
    R = 8.0
    w, bad = mask_weights.interface_weights(lambda x, y, z: R * R - x * x - y * y - z * z, dims, sp, mask)
    assert not bad
    _say("droplet256: oracle on wavelengths", [round(wl[i], 3) for i in pick])
    want = oracle.propagate(sub, lc, mask=mask, weights=w)
    _say("droplet256: FP32 graph, 294 propagations")
    e32 = _context_run(s, lc, mask, formula, precision=32, wl_batch=0, use_graph=True)
    _say("droplet256: FP64 batch, 294 propagations")
    e64 = _context_run(s, lc, mask, formula, precision=64, wl_batch=0)
    err64 = _rel_all(e64[pick], want)
    err32 = _rel_all(e32[pick], want)
    # every output of the sweep: FP32 against the FP64 path
    err32_all = _rel_all(e32, e64)
    _say("droplet256: FP64 max %.3e, FP32 max %.3e, FP32 vs FP64 over all 294 max %.3e median %.3e"
         % (err64.max(), err32.max(), err32_all.max(), np.median(err32_all)))
    assert np.isfinite(e32).all() and np.isfinite(e64).all()
    assert err64.max() < TOL64, err64
    assert err32.max() < TOL32, err32
    assert err32_all.max() < TOL32, err32_all.max()


@pytest.mark.timeout(600)
def test_cholesteric1024_full_fp64_and_fp32_against_oracle():
    _say("cholesteric1024: director")
    s, lc, _, _ = fields.workload("cholesteric1024")
    _say("cholesteric1024: oracle")
    want = oracle.propagate(s, lc)
    _say("cholesteric1024: FP64 one-shot")
    e64 = bpm_backend.propagate(json.dumps(s), lc)
    _say("cholesteric1024: FP32 graph")
    e32 = _context_run(s, lc, precision=32, wl_batch=0, use_graph=True)
    err64, err32 = _rel_all(e64, want), _rel_all(e32, want)
    _say("cholesteric1024: FP64 max %.3e, FP32 max %.3e" % (err64.max(), err32.max()))
    assert err64.max() < TOL64, err64
    assert err32.max() < TOL32, err32


def test_serving_trace_over_the_benchmark_meshes():
    """Four jobs, one per mesh size of the serving workload, with the three
    director generators, submitted together: each result equals its
    standalone run bit for bit."""
    ring = fields.ring_wavevectors(0.2, 2)
    specs = [
        {"job": 0, "arrival_s": 0.0, "mesh": (128, 128, 40), "generator": "random-smooth",
         "wavelengths": list(np.linspace(0.40, 0.80, 21)), "wavevectors": ring.tolist()},
        {"job": 1, "arrival_s": 0.0, "mesh": (256, 256, 80), "generator": "droplet",
         "wavelengths": [0.45, 0.6, 0.75], "wavevectors": ring.tolist()},
        {"job": 2, "arrival_s": 0.0, "mesh": (512, 512, 100), "generator": "twisted",
         "wavelengths": [0.5, 0.6], "wavevectors": ring[:2].tolist()},
        {"job": 3, "arrival_s": 0.0, "mesh": (1024, 1024, 200), "generator": "twisted",
         "wavelengths": [0.6], "wavevectors": ring[:1].tolist()},
    ]
    cache = {}
    jobs = []
    with serving.Server(max_active=4) as srv:
        for sp_ in specs:
            lc, mask, formula, spacing = serving.director(sp_["mesh"], sp_["generator"], cache)
            js = json.dumps(serving.job_settings(sp_, spacing))
            _say("serving: submit", sp_["mesh"], sp_["generator"])
            jobs.append((js, lc, mask, formula, srv.submit(js, lc, mask_vals=mask, mask_formula=formula or "",
                                                           precision=32, wl_batch=0)))
        infos = [j[-1].wait() for j in jobs]
        st = srv.stats()
    assert all(i["state"] == "done" for i in infos), infos
    assert st["peak_active"] >= 2, st
    for js, lc, mask, formula, job in jobs:
        _say("serving: standalone", json.loads(js)["Physics settings"]["Initial conditions"]["Mesh dimensions"])
        alone = bpm_backend.propagate(js, lc, mask_vals=mask, mask_formula=formula or "", precision=32, wl_batch=0)
        assert np.isfinite(alone).all()
        assert np.array_equal(job.E, alone)
