"""Serving loop on the GPU: isolation (each result is the same alone and in
the mix), admission by footprint under a small arena, refusal of a job that
can never fit, cancellation of a queued job, refusal of invalid jobs, and a
job that fails after its march without disturbing the job beside it.  Failure conditions are all forced
from the host (sizes, settings); no kernel is made to fail.

Tolerance of "the same alone and in the mix": bit for bit (the kernels are
deterministic and the launch shapes depend only on the job).
"""
import json

import numpy as np
import pytest

from nemaktis_amd import bpm_backend, fields, serving

pytestmark = pytest.mark.gpu


def _job(dims, seed, wl=(0.55, 0.65), nq=2, gen="random"):
    Nx, Ny, Nz = dims
    sp = (0.1, 0.1, 0.1)
    mask = formula = None
    if gen == "random":
        lc = fields.random_smooth(Nx, Ny, Nz, sp, seed=seed, corr_length=0.5)
    elif gen == "droplet":
        lc, mask, formula = fields.droplet(Nx, Ny, Nz, sp, 0.3 * Nx * sp[0])
    else:
        lc = fields.twisted_nematic(Nx, Ny, Nz)
    s = fields.make_settings(dims, sp, list(wl), wavevectors=fields.ring_wavevectors(0.2, 2)[:nq], nhost="1.51",
                             na_objective=0.4)
    return json.dumps(s), lc, mask, formula


def test_isolation_alone_and_in_the_mix():
    a = _job((64, 64, 12), seed=1)
    b = _job((96, 80, 10), seed=2, gen="droplet")
    c = _job((128, 128, 8), seed=3, wl=(0.6,), nq=1, gen="twisted")
    with serving.Server(max_active=3) as srv:
        print("[serving] job A alone", flush=True)
        alone = srv.submit(a[0], a[1], precision=32)
        assert alone.wait()["state"] == "done"
        print("[serving] mix B, A, C(FP64), A", flush=True)
        mix = [srv.submit(b[0], b[1], mask_vals=b[2], mask_formula=b[3], precision=32),
               srv.submit(a[0], a[1], precision=32),
               srv.submit(c[0], c[1], precision=64, wl_batch=1),
               srv.submit(a[0], a[1], precision=32)]
        infos = [j.wait() for j in mix]
        assert all(i["state"] == "done" for i in infos), infos
        st = srv.stats()
    assert np.array_equal(mix[1].E, alone.E)
    assert np.array_equal(mix[3].E, alone.E)
    # the masked job and the FP64 job equal the plain one-shot runs
    eb = bpm_backend.propagate(b[0], b[1], mask_vals=b[2], mask_formula=b[3], precision=32, wl_batch=0)
    ec = bpm_backend.propagate(c[0], c[1])
    assert np.array_equal(mix[0].E, eb)
    assert np.array_equal(mix[2].E, ec)
    assert st["completed"] == 5 and st["failed"] == 0
    assert st["peak_active"] >= 2, st  # the mix overlapped on the device


def test_graph_jobs_match_eager_jobs():
    """Jobs captured as graphs (their buffers carved from the arena) give the
    eager jobs' results bit for bit, alone and beside each other, with the
    coefficients on the march stream (built-in) or on a second stream."""
    a = _job((64, 64, 12), seed=4)
    b = _job((96, 80, 10), seed=5, gen="droplet")
    t = {"x_chunks": 16, "x_rows": 2, "y_chunks": 16, "y_cols": 8, "coef_serial": 0}
    with serving.Server(max_active=3) as srv:
        print("[serving] eager A, B", flush=True)
        ea = srv.submit(a[0], a[1], precision=32)
        eb = srv.submit(b[0], b[1], mask_vals=b[2], mask_formula=b[3], precision=32)
        assert ea.wait()["state"] == "done" and eb.wait()["state"] == "done"
        print("[serving] graph A, B, A (coefficients on a second stream)", flush=True)
        ga = srv.submit(a[0], a[1], precision=32, use_graph=True)
        gb = srv.submit(b[0], b[1], mask_vals=b[2], mask_formula=b[3], precision=32, use_graph=True)
        gs = srv.submit(a[0], a[1], precision=32, use_graph=True, tuning=t)
        infos = [j.wait() for j in (ga, gb, gs)]
        assert all(i["state"] == "done" for i in infos), infos
        es = srv.submit(a[0], a[1], precision=32, tuning=t)
        assert es.wait()["state"] == "done"
    assert np.array_equal(ga.E, ea.E)
    assert np.array_equal(gb.E, eb.E)
    assert np.array_equal(gs.E, es.E)


def test_admission_by_footprint():
    a = _job((64, 64, 10), seed=4)
    with serving.Server(max_active=4, hbm_budget=1 << 30) as probe:
        one = probe.submit(a[0], a[1], precision=32)
        need = one.wait()["footprint_bytes"]
    budget = 2 * need + need // 2  # room for two jobs at a time
    with serving.Server(max_active=4, hbm_budget=budget) as srv:
        jobs = [srv.submit(a[0], a[1], precision=32) for _ in range(6)]
        infos = [j.wait() for j in jobs]
        st = srv.stats()
    assert all(i["state"] == "done" for i in infos)
    assert st["peak_active"] <= 2, st
    assert st["arena_peak_bytes"] <= budget
    for j in jobs[1:]:
        assert np.array_equal(j.E, jobs[0].E)


def test_job_larger_than_the_arena_is_refused():
    a = _job((64, 64, 10), seed=5)
    with serving.Server(max_active=2, hbm_budget=8 << 20) as srv:
        with pytest.raises(serving.ServingError, match="arena holds"):
            srv.submit(a[0], a[1], precision=64)
        assert srv.stats()["submitted"] == 0


def test_cancel_queued_job():
    big = _job((256, 256, 30), seed=6, wl=(0.5, 0.6, 0.7), nq=3)
    small = _job((64, 64, 8), seed=7)
    with serving.Server(max_active=1) as srv:
        first = srv.submit(big[0], big[1], precision=64, wl_batch=1)
        second = srv.submit(small[0], small[1], precision=32)
        cancelled = second.cancel()
        i1, i2 = first.wait(), second.wait()
    if cancelled:
        assert i2["state"] == "cancelled"
        assert not second.E.any()  # never written
    else:  # the first job had already finished: the second ran normally
        assert i2["state"] == "done"
    assert i1["state"] == "done"


def test_invalid_jobs_are_refused_and_leave_the_others_alone():
    good = _job((64, 64, 8), seed=8)
    # 4099 points per line: beyond the line solvers (at most 4097 points per
    # axis), refused at submission like malformed settings
    bad = _job((4099, 4, 3), seed=9)
    with serving.Server(max_active=2) as srv:
        with pytest.raises(serving.ServingError):
            srv.submit("{not json", good[1])
        with pytest.raises(serving.ServingError, match="4097"):
            srv.submit(bad[0], bad[1], precision=32)
        j_good = srv.submit(good[0], good[1], precision=32)
        ig = j_good.wait()
        assert srv.stats()["submitted"] == 1
    assert ig["state"] == "done"
    ref = bpm_backend.propagate(good[0], good[1], precision=32, wl_batch=0)
    assert np.array_equal(j_good.E, ref)


def test_short_poisson_trace():
    specs = serving.poisson_trace(6, rate_hz=20.0, seed=11, meshes=[(64, 64, 10), (128, 128, 12)], max_wl=3,
                                  max_q=2)
    with serving.Server(max_active=3) as srv:
        res = serving.run_trace(srv, specs)
    assert res["jobs"] == 6
    assert res["p99_ms"] >= res["p50_ms"] > 0
    assert res["updates_per_s"] > 0


def test_job_failing_after_its_march_leaves_its_neighbour_intact(tmp_path):
    """A job whose bulk output cannot be written fails after its kernels ran
    (its arena range is handed back only once its stream has drained); the job
    beside it equals its standalone run bit for bit."""
    good = _job((96, 96, 12), seed=12)
    bad_s = json.loads(_job((64, 64, 10), seed=13)[0])
    blocker = tmp_path / "not_a_folder"
    blocker.write_text("a file where the results folder should be")
    bad_s["Algorithm settings"]["General"]["Results folder name"] = str(blocker)
    bad_s["Postprocessor settings"]["Bulk output"] = {"Activate": True, "Base name": "x"}
    bad_lc = fields.random_smooth(64, 64, 10, (0.1, 0.1, 0.1), seed=13, corr_length=0.5)
    with serving.Server(max_active=2) as srv:
        print("[serving] failing bulk job beside a good job", flush=True)
        jb = srv.submit(json.dumps(bad_s), bad_lc, precision=32)
        jg = srv.submit(good[0], good[1], precision=32)
        ib, ig = jb.wait(), jg.wait()
        jg2 = srv.submit(good[0], good[1], precision=32)
        ig2 = jg2.wait()
    assert ib["state"] == "failed", ib
    assert ig["state"] == "done" and ig2["state"] == "done"
    ref = bpm_backend.propagate(good[0], good[1], precision=32, wl_batch=0)
    assert np.array_equal(jg.E, ref) and np.array_equal(jg2.E, ref)


def test_mask_formula_mismatch_notice_in_the_job_record():
    """A job whose mask formula disagrees with its mask: the reference's notice
    (UtilitiesLib/VectorField.cpp:107-111, found by the device weights) in the
    job record, and the default weights (the same output as no formula)."""
    dims, sp = (48, 40, 8), (0.2, 0.2, 0.2)
    lc, mask, _ = fields.droplet(*dims, sp, 2.5)
    s = json.dumps(fields.make_settings(dims, sp, [0.6], nhost="1.51", na_objective=0.4))
    with serving.Server(max_active=2) as srv:
        print("[serving] mismatching mask formula", flush=True)
        j = srv.submit(s, lc, mask_vals=mask, mask_formula="0.05-z", precision=32)
        info = j.wait()
    assert info["state"] == "done", info
    assert info["notice"] == "Mask values do not corresponds to the given formula, will ignore interpolation correction"
    ref = bpm_backend.propagate(s, lc, mask_vals=mask, mask_formula="", precision=32, wl_batch=0)
    assert np.array_equal(j.E, ref)
