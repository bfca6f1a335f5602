#!/usr/bin/env python
"""Benchmark of the bpm-solver propagation step on one MI355X.

Workload (default): 512x512x100 mesh, 0.1 um spacing, seeded smoothed random
director (1 um correlation), no = 1.52, ne = 1.74, wavelength 0.55 um, normal
incidence, 2 input polarisations.  One step = one full propagation of that
job (permittivity and coefficient precompute, input/output Fresnel, the
99-slab z-march and the FFT screen filter) with the director already
resident in HBM.  `value` is the FP64 path (complex128 march, the reference's
arithmetic, one wavelength at a time, z-march replayed as a hipGraph); beside
it: the same path with eager launches, the FP32 batched path (complex64 march,
FP64 coefficients) with its error against FP64, the roofline of the dominant
kernel of each, the FP64 CPU oracle on the same workload (median of 3), the
one-shot drop-in call timed end to end, and the serving, imaging, bulk-output
and droplet-sweep legs.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FAMILIES = ["material", "coefs", "boundary", "reserved", "explicit", "solve_x", "solve_y", "output", "filter", "setup"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


KERNELS = {"solve_x": "k_xsolve", "solve_y": "k_ysolve", "boundary": "k_boundary", "coefs": "k_slab_coefs",
           "explicit": "k_explicit"}


def algorithmic_bytes(fam, c, nprops, nslots, NP, Nfull, fd, nw):
    """Unique HBM bytes one launch of a kernel family must move: every array
    element it reads or writes counted once (c = bytes per complex value,
    NP march-grid points, Nfull full-grid points).
      line solves, first launch of a half-step: per propagation the source
      field (2 components) in, the solution (2 components) out.  Neumann
      launch i of nw: source in, the sum so far in (i > 0; launch 0 adds the
      source itself), the sum out, the solution out for the next launch
      (i < nw - 1).  Per wavelength slot the coefficient planes of the taps
      (complex planes count 1, imaginary-only planes 1/2) and the 6
      line-matrix planes: x first 15 + 6, y first 13 + 6, Neumann 9 + 6.
      The family average weighs one first launch and the nw Neumann
      launches."""
    if fam == "boundary":
        return 4 * nprops * NP * c + 6 * nslots * NP * c
    if fam == "explicit":
        # (I + mu A_y) or (I + mu A_x) of both components: 15 or 13 coefficient planes, averaged
        return 4 * nprops * NP * c + 14 * nslots * NP * c
    if fam in ("solve_x", "solve_y"):
        total = (4 * nprops + (21 if fam == "solve_x" else 19) * nslots) * NP * c
        for i in range(nw):
            planes = 2 + (2 if i > 0 else 0) + 2 + (2 if i + 1 < nw else 0)
            total += (planes * nprops + 15 * nslots) * NP * c
        return total / (1 + nw)
    if fam == "coefs":
        # director planes iz, iz+1 in (FP64); per slot 16 complex + 24 imaginary-only coefficient planes,
        # 3 phase-operator planes and 6 line-matrix planes out
        return 2 * fd * 8 * Nfull + 37 * nslots * NP * c
    return None


TRAFFIC_FILE_IMAGING = os.path.join(ROOT, "profiles", "hbm_traffic_imaging_droplet256.txt")


def measured_traffic(kernel, path):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy); (None, None) when there is none."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == kernel and parts[1] == "hbm_bytes_per_launch":
                    return float(parts[2]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def probe(kind, steps):
    """Short runs for hardware-counter passes (rocprofv3 --pmc ... -- python
    bench.py --probe KIND): 'copy' copies 256 MiB device to device `steps`
    times (counter calibration), 'march' runs the FP32 workload, 'march-droplet'
    the 256x256x80 droplet sweep (294 propagations, FP32, built-in shapes),
    'march64' the FP64 path of the workload (one wavelength at a time),
    'imaging' the refocus + micrograph step, all with eager launches."""
    if kind == "copy":
        import torch
        n = 64 << 20
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(steps):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: %d copies of %d bytes" % (steps, 4 * n))
        return
    if kind == "imaging":
        of, cmf = imaging_workload()
        from nemaktis_amd import viewer
        fv = viewer.FieldViewer(of, cmf=cmf)
        fv.NA_condenser = IMAGING["NA_condenser"]
        for _ in range(steps):
            of.focus_fields(IMAGING["z_focus"])
            fv.render()
        log("imaging probe: %d refocus + render steps, last %s" % (steps, of.timings()))
        of.close()
        return
    from nemaktis_amd import fields, bpm_backend
    tuning = None  # built-in launch shapes, as the benchmark's legs
    precision, wl_batch = 32, 0
    if kind == "march-droplet":
        s, lc, mask, formula = fields.workload("droplet256")
    else:
        s, lc, mask, formula = fields.workload("random512")
        if kind == "march64":
            precision, wl_batch = 64, 1
    js = json.dumps(s)
    n_E = int(np.prod(fields.output_shape(s)))
    with bpm_backend.Context(js, lc, n_E, mask_vals=mask, mask_formula=formula or "", precision=precision,
                             wl_batch=wl_batch, use_graph=False, tuning=tuning) as ctx:
        for _ in range(steps):
            st = ctx.propagate()
    log("march probe: %d propagations, last %.2f ms, tuning %s" % (steps, st["total_ms"], tuning))


def serve_leg(n_jobs, rate_hz, seed, max_active):
    """Seeded Poisson trace of micrograph jobs against the serving loop."""
    from nemaktis_amd import serving
    specs = serving.poisson_trace(n_jobs, rate_hz, seed)
    cache = {}
    t0 = time.perf_counter()
    for sp in specs:
        serving.director(sp["mesh"], sp["generator"], cache)
    log("serving trace: %d jobs, directors generated in %.1f s" % (n_jobs, time.perf_counter() - t0))
    with serving.Server(max_active=max_active) as srv:
        res = serving.run_trace(srv, specs, cache, precision=32, wl_batch=0, log=log)
    res.update({"trace": "%d jobs, Poisson arrivals %.2f/s, seed %d, meshes 128^2x40 / 256^2x80 / 512^2x100 / "
                         "1024^2x200, 1-21 wavelengths x 1-7 wavevectors x 2 polarisations, complex64"
                         % (n_jobs, rate_hz, seed), "max_active": max_active})
    return res


IMAGING = {"Nx": 256, "Ny": 256, "L": 31.875, "wavelengths": list(np.linspace(0.40, 0.80, 21)), "Nr": 2,
           "NA_condenser": 0.2, "NA_objective": 0.4, "z_focus": 1.0}


def imaging_bytes(kernel, Nl, Nq, N, nq_used):
    """Unique HBM bytes of one launch of an imaging kernel (complex128 = 16 B):
      filter   writes one plane per (wavelength, wavevector)
      rows     read and write the 4 field planes of every (wavelength, wavevector)
      columns  the same plus one read of the filter plane
      spectrum reads the 4 planes of every used (wavelength, wavevector), writes Nl planes of float64
      colour   reads Nl float64 planes, writes the spectrum (Nl) and image (3) float64 planes"""
    return {"filter": 16 * Nl * Nq * N, "rows_fwd": 128 * Nl * Nq * N, "rows_bwd": 128 * Nl * Nq * N,
            "columns": 144 * Nl * Nq * N, "spectrum": 64 * Nl * nq_used * N + 8 * Nl * N,
            "colour": 16 * Nl * N + 24 * N}[kernel]


def imaging_workload(seed=7):
    from nemaktis_amd import viewer
    w = IMAGING
    of = viewer.OpticalFields(wavelengths=w["wavelengths"], max_NA_objective=w["NA_objective"],
                              max_NA_condenser=w["NA_condenser"], N_radial_wavevectors=w["Nr"],
                              mesh_lengths=(w["L"], w["L"]), mesh_dimensions=(w["Nx"], w["Ny"]))
    rng = np.random.default_rng(seed)
    v = of.vals
    v.real[...] = rng.standard_normal(v.shape)
    v.imag[...] = rng.standard_normal(v.shape)
    Nl = len(w["wavelengths"])
    cmf = np.stack([np.exp(-((np.array(w["wavelengths"]) - c) / 0.05) ** 2) for c in (0.6, 0.55, 0.45)], -1)
    return of, cmf / cmf[:, 1].sum() * 2


def imaging_leg(steps, warmup, with_cpu):
    """Objective refocus + micrograph synthesis of a 21-wavelength x 7-wavevector 256^2 output on the device
    (fields resident in HBM; only the device work is timed)."""
    from nemaktis_amd import viewer
    w = IMAGING
    of, cmf = imaging_workload()
    fv = viewer.FieldViewer(of, cmf=cmf)
    fv.NA_condenser = w["NA_condenser"]
    fv.lower_waveplate = "Tint-sensitive"
    fv.lower_waveplate_angle = 45
    for _ in range(warmup):
        of.focus_fields(w["z_focus"])
        fv.render()
    per = {k: 0.0 for k in ("filter", "rows_fwd", "columns", "rows_bwd", "spectrum", "colour")}
    t0 = time.perf_counter()
    for _ in range(steps):
        of.focus_fields(w["z_focus"])
        fv.render()
        for k, v in of.timings().items():
            per[k] += v
    wall = time.perf_counter() - t0
    Nl, Nq, N = len(w["wavelengths"]), 7, w["Nx"] * w["Ny"]
    avg = {k: v / steps for k, v in per.items()}
    dom = max(avg, key=avg.get)
    algo = imaging_bytes(dom, Nl, Nq, N, Nq)
    achieved = algo / (avg[dom] * 1e-3) / 1e9
    kname = {"columns": "k_view_columns", "rows_fwd": "k_fft_rows", "rows_bwd": "k_fft_rows",
             "spectrum": "k_view_spectrum", "filter": "k_view_filter", "colour": "k_view_colour"}[dom]
    traffic, traffic_src = measured_traffic(kname, TRAFFIC_FILE_IMAGING)
    out = {"metric": "micrographs/s (refocus + 21-wavelength colour micrograph)", "value": steps / wall,
           "unit": "micrographs/s", "ms_per_step": 1e3 * wall / steps,
           "device_ms": {k: round(v, 4) for k, v in avg.items()},
           "workload": "%d wavelengths x %d wavevectors x 4 planes of %dx%d complex128 (the droplet256 output "
                       "shape), z_focus %.1f, NA %.1f / condenser %.1f, tint plate, crossed polarisers"
                       % (Nl, Nq, w["Nx"], w["Ny"], w["z_focus"], w["NA_objective"], w["NA_condenser"]),
           "data": "synthetic (seeded normal fields)",
           "roofline": {"bound": "hbm", "kernel": kname, "avg_launch_ms": avg[dom], "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic, "traffic_source": traffic_src, "algorithmic_bytes_per_launch": algo,
                        "all": {k: round(imaging_bytes(k, Nl, Nq, N, Nq) / (avg[k] * 1e-3) / 1e9, 1)
                                for k in avg if avg[k] > 0}}}
    of.close()
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import viewer_oracle as vo
        nl = 3
        vals = of.vals[:nl]
        d = w["L"] / (w["Nx"] - 1)
        t0 = time.perf_counter()
        foc = vo.focus(vals, d, d, w["wavelengths"][:nl], 1.0, of.get_wavevectors(), w["NA_objective"],
                       w["z_focus"])
        spec = vo.spectrum(foc, w["wavelengths"][:nl], dict(lower_waveplate="Tint-sensitive",
                                                             lower_waveplate_angle=45,
                                                             NA_condenser=w["NA_condenser"]),
                           w["Nr"], w["NA_condenser"])
        vo.colour(spec, cmf[:nl])
        dt = (time.perf_counter() - t0) * Nl / nl
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "micrographs/s", "cores": 1, "kind": "port",
                               "sample": "numpy FP64 oracle on %d of the %d wavelengths (all 7 wavevectors), "
                                         "scaled by %d/%d" % (nl, Nl, Nl, nl)}
        out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def bulk_leg(settings, lc, tuning, folder):
    """Propagation of the benchmark job with "Bulk output" active (every plane kept in HBM, then written as
    a VTI file) against the same job without it: device ms, file size and write time."""
    from nemaktis_amd import bpm_backend, fields
    s = json.loads(json.dumps(settings))
    s["Algorithm settings"]["General"]["Results folder name"] = folder
    s["Postprocessor settings"]["Bulk output"] = {"Activate": True, "Base name": "bench_bulk"}
    n_E = int(np.prod(fields.output_shape(s)))
    out = {}
    for tag, ss in (("without", settings), ("with", s)):
        with bpm_backend.Context(json.dumps(ss), lc, n_E, precision=32, wl_batch=0, use_graph=True,
                                 tuning=tuning) as ctx:
            ctx.propagate()
            t0 = time.perf_counter()
            st = ctx.propagate()
            wall = time.perf_counter() - t0
        out[tag] = {"device_ms": st["total_ms"], "wall_ms": 1e3 * wall, "footprint_gb": st["footprint_bytes"] / 1e9}
    path = os.path.join(folder, "bench_bulk.vti")
    size = os.path.getsize(path)
    os.remove(path)
    out["file_gb"] = size / 1e9
    out["write_s"] = (out["with"]["wall_ms"] - out["with"]["device_ms"]) / 1e3
    out["write_gbs"] = out["file_gb"] / max(1e-9, out["write_s"])
    out["device_overhead"] = out["with"]["device_ms"] / out["without"]["device_ms"] - 1
    return out


def rel_errors(a, b):
    a = a.reshape(a.shape[0], a.shape[1], 2, -1)
    b = b.reshape(b.shape[0], b.shape[1], 2, -1)
    e = np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)
    return e.reshape(-1)


def droplet_leg(steps, warmup):
    """The 256x256x80 droplet sweep (21 wavelengths x 7 wavevectors x 2
    polarisations = 294 propagations marched together, mask and mask formula),
    FP32 with graph replay and the built-in launch shapes: wall time per step
    with the director resident in HBM."""
    from nemaktis_amd import bpm_backend, fields
    s, lc, mask, formula = fields.workload("droplet256")
    n_E = int(np.prod(fields.output_shape(s)))
    with bpm_backend.Context(json.dumps(s), lc, n_E, mask_vals=mask, mask_formula=formula or "", precision=32,
                             wl_batch=0, use_graph=True) as ctx:
        for _ in range(warmup):
            ctx.propagate()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.propagate()
        ms = (time.perf_counter() - t0) * 1e3 / steps
    updates = fields.updates_per_run(s)
    out = {"workload": "droplet 256x256x80, 21 wavelengths x 7 wavevectors x 2 polarisations, FP32, graph, "
                       "built-in shapes", "ms_per_step": ms, "value": updates / ms * 1e3,
           "unit": "voxel-slice updates/s", "steps": steps, "warmup": warmup}
    # roofline of the batch's dominant march kernel: HIP-event launch times of an eager pass, the same
    # per-update algorithmic bytes as the headline (64 B of FP32 field traffic + 24 B of director per slot,
    # shared by the slot's 2 Nq propagations) times the updates one launch advances
    with bpm_backend.Context(json.dumps(s), lc, n_E, mask_vals=mask, mask_formula=formula or "", precision=32,
                             wl_batch=0, profile=True) as ctx:
        ctx.propagate()
        st = ctx.propagate()
    fam_ms, fam_n = st["family_ms"], st["family_launches"]
    march = [f for f in FAMILIES if f not in ("setup", "output", "filter") and fam_n.get(f)]
    dom = max(march, key=lambda f: fam_ms[f])
    avg_ms = fam_ms[dom] / fam_n[dom]
    co = s["Physics settings"]["Coefficients"]
    nq = len(co["Wavevectors"]) // 2
    Nx, Ny, _ = s["Physics settings"]["Initial conditions"]["Mesh dimensions"]
    nprops = len(co["Wavelengths"]) * nq * 2
    per_update = 64 + 24.0 / (2 * nq)
    algo = per_update * nprops * Nx * Ny
    achieved = algo / (avg_ms * 1e-3) / 1e9
    traffic, src = measured_traffic(KERNELS.get(dom, dom),
                                    os.path.join(ROOT, "profiles", "hbm_traffic_droplet256_fp32.txt"))
    out["roofline"] = {"bound": "hbm", "kernel": KERNELS.get(dom, dom), "family": dom,
                       "avg_launch_ms": avg_ms, "launches": fam_n[dom], "achieved": round(achieved, 2),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                       "traffic": traffic, "traffic_source": src, "algorithmic_bytes_per_launch": algo,
                       "algorithmic_basis": "%.1f B per voxel-slice update x %d updates per launch "
                                            "(%d propagations x %dx%d)" % (per_update, nprops * Nx * Ny, nprops,
                                                                           Nx, Ny),
                       "step_frac": round(updates * per_update / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
    return out


def run_gpu_leg(js, lc, n_E, shape, precision, wl_batch, use_graph, steps, warmup, tuning=None):
    from nemaktis_amd import bpm_backend
    with bpm_backend.Context(js, lc, n_E, precision=precision, wl_batch=wl_batch, use_graph=use_graph,
                             tuning=tuning) as ctx:
        for _ in range(warmup):
            ctx.propagate()
        dev_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            st = ctx.propagate()  # returns once the device work of the step has completed
            dev_ms.append(st["total_ms"])
        wall = time.perf_counter() - t0
        out = ctx.download(np.empty(shape, np.complex128))
        return wall, dev_ms, st, out


def run_profile_leg(js, lc, n_E, precision, wl_batch, steps, tuning=None):
    from nemaktis_amd import bpm_backend
    fam_ms = {f: 0.0 for f in FAMILIES}
    fam_n = {f: 0 for f in FAMILIES}
    with bpm_backend.Context(js, lc, n_E, precision=precision, wl_batch=wl_batch, profile=True,
                             tuning=tuning) as ctx:
        ctx.propagate()  # warm-up
        for _ in range(steps):
            st = ctx.propagate()
            for f in FAMILIES:
                fam_ms[f] += st["family_ms"][f]
                fam_n[f] += st["family_launches"][f]
    return fam_ms, fam_n


def host_facts():
    """CPU model, hardware threads available to this process (affinity and
    OMP_NUM_THREADS) and memory of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    mem_gb = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    mem_gb = round(int(line.split()[1]) / 1024 ** 2, 1)
                    break
    except OSError:
        pass
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"cpu_model": model, "threads": threads, "memory_gb": mem_gb}


def cpu_baseline(settings_fn, lc_full, nz_sample, repeats=3):
    """The FP64 oracle (the repository's CPU restatement, OpenMP on every
    available host thread) on the same director: median of `repeats` runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from nemaktis_amd import fields
    s = settings_fn(nz_sample)
    lc = np.ascontiguousarray(lc_full[:nz_sample])
    facts = host_facts()
    times, out = [], None
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = oracle.propagate(s, lc, nthreads=facts["threads"])
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return fields.updates_per_run(s) / dt, facts, times, out


def drop_in_leg(js, lc, shape, precision):
    """The one-shot backend call as nemaktis makes it (host arrays in, host
    array out), timed end to end, beside its parts on a context: creation
    (allocation and director upload), the device work, the download."""
    from nemaktis_amd import bpm_backend
    E = np.zeros(shape, np.complex128)
    st = {}
    bpm_backend.run_backend_without_mask(js, lc, E, precision=precision, strict=True)  # first call loads the code
    t0 = time.perf_counter()
    bpm_backend.run_backend_without_mask(js, lc, E, precision=precision, strict=True, stats=st)
    wall = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx = bpm_backend.Context(js, lc, E.size, precision=precision)
    create = time.perf_counter() - t0
    try:
        t0 = time.perf_counter()
        dev = ctx.propagate()
        prop = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.download(E)
        down = time.perf_counter() - t0
    finally:
        ctx.close()
    return {"call_ms": 1e3 * wall, "device_ms": st.get("total_ms"),
            "context_create_ms": 1e3 * create, "director_gb": lc.nbytes / 1e9,
            "upload_gbs": lc.nbytes / 1e9 / create, "propagate_wall_ms": 1e3 * prop,
            "propagate_device_ms": dev["total_ms"], "download_ms": 1e3 * down,
            "download_gbs": E.nbytes / 1e9 / max(down, 1e-9),
            "note": "one-shot run_backend_without_mask on host arrays (copy-in of the director, eager launches, "
                    "copy-out of E); context_create = allocation + director upload"}


def drop_in_masked_leg():
    """The masked one-shot call on the droplet sweep (run_backend_with_mask
    with the mask formula, FP32, all wavelengths at once), host arrays in and
    out, timed end to end beside its device time; the call uploads the director
    and the LC flags and computes the interface weights on the device."""
    from nemaktis_amd import bpm_backend, fields
    s, lc, mask, formula = fields.workload("droplet256")
    js = json.dumps(s)
    E = np.zeros(fields.output_shape(s), np.complex128)
    bpm_backend.run_backend_with_mask(js, formula, lc, mask, E, precision=32, wl_batch=0, strict=True)
    st = {}
    t0 = time.perf_counter()
    bpm_backend.run_backend_with_mask(js, formula, lc, mask, E, precision=32, wl_batch=0, strict=True, stats=st)
    wall = time.perf_counter() - t0
    return {"call_ms": 1e3 * wall, "device_ms": st.get("total_ms"), "director_gb": lc.nbytes / 1e9,
            "mask_gb": mask.nbytes / 1e9, "output_gb": E.nbytes / 1e9,
            "note": "one-shot run_backend_with_mask, droplet 256x256x80 with its mask formula, 294 propagations, "
                    "FP32, eager launches"}


def roofline_of(js, lc, n_E, precision, wl_batch, steps, Nx, Ny, nprops, nq, fd, nw, tuning=None):
    """Dominant march kernel family of the path, its average launch time
    (HIP events on its launch stream, eager), and its bandwidth two ways:
    the workload's algorithmic bytes per voxel-slice update (64 / 128 B of
    field traffic + 24 B of director per slot, shared by the slot's 2 Nq
    propagations) times the updates one launch advances (every propagation
    at every point of the full grid: nprops * Nx * Ny), and the kernel's own
    unique bytes per launch (its field, coefficient and line-row planes)."""
    fam_ms, fam_n = run_profile_leg(js, lc, n_E, precision, wl_batch, steps, tuning)
    march_fams = [f for f in FAMILIES if f not in ("setup", "output", "filter") and fam_n[f]]
    dom = max(march_fams, key=lambda f: fam_ms[f])
    avg_ms = fam_ms[dom] / fam_n[dom]
    per_update = (64 if precision == 32 else 128) + 24.0 / (2 * nq)
    units = nprops * Nx * Ny
    algo = per_update * units
    achieved = algo / (avg_ms * 1e-3) / 1e9
    c = 8 if precision == 32 else 16
    local = algorithmic_bytes(dom, c, nprops, nprops // (2 * nq), (Nx - 1) * (Ny - 1), Nx * Ny, fd, nw)
    tag = "fp32" if precision == 32 else "fp64"
    traffic, traffic_src = measured_traffic(KERNELS.get(dom, dom), os.path.join(ROOT, "profiles",
                                                                              "hbm_traffic_random512_%s.txt" % tag))
    return {"bound": "hbm", "kernel": KERNELS.get(dom, dom), "family": dom, "precision": precision,
            "avg_launch_ms": avg_ms, "launches": fam_n[dom], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": algo,
            "algorithmic_basis": "%.1f B per voxel-slice update x %d updates per launch (%d propagations x %dx%d)"
                                 % (per_update, units, nprops, Nx, Ny),
            "kernel_local_bytes_per_launch": local,
            "kernel_local_frac": round(local / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if local else None,
            "family_share": {f: round(fam_ms[f] / max(1e-9, sum(fam_ms.values())), 4) for f in FAMILIES}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-nz", type=int, default=100, help="planes of the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--tune", action="store_true", help="autotuned line-solver shapes instead of the built-in ones")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop trace")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=0.5, help="mean job arrivals per second")
    ap.add_argument("--serve-only", action="store_true", help="run the serving trace alone (one JSON line)")
    ap.add_argument("--serve-active", type=int, default=4, help="jobs on the device at once")
    ap.add_argument("--no-imaging", action="store_true", help="skip the refocus + micrograph leg")
    ap.add_argument("--no-bulk", action="store_true", help="skip the bulk-output leg")
    ap.add_argument("--no-droplet", action="store_true", help="skip the 294-propagation droplet sweep leg")
    ap.add_argument("--no-dropin", action="store_true", help="skip the one-shot drop-in call leg")
    ap.add_argument("--bulk-dir", default="/tmp/nemaktis_amd_bench", help="folder of the bulk-output file")
    ap.add_argument("--probe", choices=["copy", "march", "march64", "march-droplet", "imaging"],
                    help="short run for a hardware-counter pass")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the bpm march shards by independent (wavelength, wavevector, "
              "polarisation) propagations and this build runs on a single MI355X; --gpus %d is not supported"
              % max(args.gpus, world), file=sys.stderr)
        sys.exit(2)

    if args.probe:
        probe(args.probe, args.steps)
        return
    from nemaktis_amd import _lib, fields
    _lib.load()
    if args.serve_only:
        print(json.dumps({"serving": serve_leg(args.serve_jobs, args.serve_rate, 5, args.serve_active)}), flush=True)
        return
    Nx = Ny = 512
    Nz = 100
    sp = (0.1, 0.1, 0.1)
    log("generating the seeded 512x512x100 director field ...")
    lc = fields.random_smooth(Nx, Ny, Nz, sp, seed=2, corr_length=1.0)

    def settings_for(nz):
        return fields.make_settings((Nx, Ny, nz), sp, [0.55], no="1.52", ne="1.74", na_objective=0.4)

    s = settings_for(Nz)
    js = json.dumps(s)
    shape = fields.output_shape(s)
    n_E = int(np.prod(shape))
    updates = fields.updates_per_run(s)

    tuning32 = tuning64 = None
    if args.tune:
        from nemaktis_amd import autotune
        log("autotuning the line-solver shapes (shallow copies of the job) ...")
        tuning64, _ = autotune.tune(s, lc, precision=64, wl_batch=1, use_cache=True, log=log, use_graph=True)
        tuning32, _ = autotune.tune(s, lc, precision=32, wl_batch=0, use_cache=True, log=log)

    # headline: the FP64 path (complex128 march, the reference's arithmetic),
    # one wavelength at a time, z-march replayed as a hipGraph
    log("FP64 path (graph) ...")
    wall64, dev64, st64, out64 = run_gpu_leg(js, lc, n_E, shape, 64, 1, True, args.steps, args.warmup, tuning64)
    ms64 = 1e3 * wall64 / args.steps
    value = updates * args.steps / wall64

    log("FP64 plain path (eager launches) ...")
    k64e = max(1, args.steps // 2)
    walle, deve, _, oute = run_gpu_leg(js, lc, n_E, shape, 64, 1, False, k64e, 1, tuning64)
    fp64_eager = {"value": updates * k64e / walle, "ms_per_step": 1e3 * walle / k64e, "steps": k64e,
                  "bit_equal_to_graph": bool(np.array_equal(oute, out64))}

    fp32 = None
    if not args.no_fp32:
        log("FP32 batched path (graph) ...")
        wall32, dev32, _, out32 = run_gpu_leg(js, lc, n_E, shape, 32, 0, True, args.steps, args.warmup, tuning32)
        errs = rel_errors(out32, out64)
        fp32 = {"value": updates * args.steps / wall32, "ms_per_step": 1e3 * wall32 / args.steps,
                "device_ms_per_step_median": float(np.median(dev32)),
                "speedup_vs_fp64_graph": ms64 / (1e3 * wall32 / args.steps),
                "speedup_vs_fp64_eager": fp64_eager["ms_per_step"] / (1e3 * wall32 / args.steps),
                "rel_l2_vs_fp64_max": float(errs.max()), "rel_l2_vs_fp64_median": float(np.median(errs)),
                "dtype": "c64 (FP64 permittivity, coefficient precompute and phase bookkeeping)"}

    log("profiled passes (kernel-family event timing) ...")
    prof_steps = max(1, min(args.steps, 3))
    roofline = roofline_of(js, lc, n_E, 64, 1, prof_steps, Nx, Ny, 2, 1, 3, 2, tuning64)
    roofline["step_achieved"] = round(updates * (128 + 12.0) / (ms64 * 1e-3) / 1e9, 2)
    roofline["step_frac"] = round(roofline["step_achieved"] / HBM_PEAK_GBS, 5)
    if fp32:
        r32 = roofline_of(js, lc, n_E, 32, 0, prof_steps, Nx, Ny, 2, 1, 3, 2, tuning32)
        r32["step_achieved"] = round(updates * (64 + 12.0) / (fp32["ms_per_step"] * 1e-3) / 1e9, 2)
        r32["step_frac"] = round(r32["step_achieved"] / HBM_PEAK_GBS, 5)
        fp32["roofline"] = r32

    cpu = None
    if not args.no_cpu:
        log("CPU oracle (median of 3) ...")
        v, facts, times, cpu_out = cpu_baseline(settings_for, lc, args.cpu_nz)
        cpu = {"value": v, "unit": "voxel-slice updates/s", "cores": facts["threads"], "kind": "port",
               "sample": "FP64 oracle on the same workload, 512x512x%d (%d slabs, 2 polarisations), median of %d "
                         "runs (%s s)" % (args.cpu_nz, args.cpu_nz - 1, len(times),
                                          ", ".join("%.2f" % t for t in times)),
               "cpu_model": facts["cpu_model"], "memory_gb": facts["memory_gb"], "build": "g++ -O3 -fopenmp"}
        if args.cpu_nz == Nz:
            cpu["fp64_gpu_rel_l2_vs_oracle_max"] = float(rel_errors(out64, cpu_out).max())

    def side_leg(name, fn):
        """a leg beside the headline: an error in it is reported in its place
        instead of losing the whole line"""
        log(name + " ...")
        try:
            return fn()
        except Exception as e:  # noqa: BLE001
            import traceback
            traceback.print_exc()
            return {"error": "%s: %s" % (type(e).__name__, e)}

    dropin = None if args.no_dropin else side_leg("one-shot drop-in call", lambda: {
        "fp64": drop_in_leg(js, lc, shape, 64), "fp32_droplet_masked": drop_in_masked_leg()})
    imaging = None if args.no_imaging else side_leg(
        "refocus + micrograph leg", lambda: imaging_leg(args.steps, args.warmup, not args.no_cpu))
    bulk = None if args.no_bulk else side_leg("bulk-output leg", lambda: bulk_leg(s, lc, tuning32, args.bulk_dir))
    droplet = None if args.no_droplet else side_leg("droplet sweep leg", lambda: droplet_leg(3, 1))
    serve = None if args.no_serve else side_leg(
        "serving loop trace", lambda: serve_leg(args.serve_jobs, args.serve_rate, 5, args.serve_active))

    line = {
        "metric": "voxel-slice updates/s at 512x512x100 (FP64 path; FP32 batched path beside it); "
                  "p50/p99 micrograph latency of the serving trace",
        "value": value, "unit": "voxel-slice updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128 (f64)", "data": "synthetic",
        "config": {"workload": "random-smooth director 512x512x100, dx=0.1um, lambda=0.55um, no=1.52, "
                               "ne=1.74, normal incidence, 2 polarisations", "updates_per_step": updates,
                   "path": "FP64, one wavelength at a time, z-march replayed as a hipGraph, built-in launch shapes"
                           if tuning64 is None else "FP64, graph, autotuned shapes",
                   "device_ms_per_step_median": float(np.median(dev64)), "line_solver_shapes": tuning64},
        "fp64_eager": fp64_eager,
        "fp32_batched": fp32,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "drop_in_call": dropin,
        "serving": serve,
        "imaging": imaging,
        "bulk_output": bulk,
        "droplet_sweep": droplet,
    }
    if cpu:
        line["vs_cpu"] = value / cpu["value"]
        if fp32:
            fp32["vs_cpu"] = fp32["value"] / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
