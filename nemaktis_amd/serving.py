"""Serving loop of the bpm-solver propagation step (Python side of the
``nmk_srv_*`` C ABI in include/nemaktis_bpm.h).

The reference backend runs one blocking propagation per call
(BeamPropagationSolver/src/PythonBackend/bpm_backend.cpp:7-164).  A
:class:`Server` instead owns one GPU and a pre-allocated HBM arena: submitted
micrograph jobs are validated at once, queued, admitted in arrival order as
soon as an execution slot and an arena range of the job's footprint are free
(a job that does not fit may be overtaken a bounded number of times), and the
admitted jobs run concurrently, each on its own HIP stream with its state kept
on the device until it finishes.

:func:`poisson_trace` and :func:`run_trace` drive a server with seeded
Poisson arrivals of jobs of uneven mesh size and report p50/p99 latency and
aggregate voxel-slice updates/s.
"""
import ctypes
import time

import numpy as np

from . import _lib, fields

__all__ = ["Server", "Job", "ServingError", "poisson_trace", "run_trace", "percentiles", "MESHES", "GENERATORS"]


class ServingError(RuntimeError):
    pass


class Job:
    """Handle of a submitted job; ``E`` receives the output fields."""

    def __init__(self, server, jid, E, inputs):
        self.server, self.id, self.E = server, jid, E
        self._inputs = inputs  # kept alive until the job has finished
        self.info = None

    def wait(self, timeout=None):
        """Waits for the job to finish and returns its info dict
        (state 'done', 'failed' or 'cancelled')."""
        info = _lib.SrvJobInfo()
        r = self.server.lib.nmk_srv_wait(self.server.h, self.id, -1.0 if timeout is None else 1e3 * timeout,
                                         ctypes.byref(info))
        if r < 0:
            raise ServingError("unknown job %d" % self.id)
        d = info.as_dict()
        if r == 1:
            raise TimeoutError("job %d still %s" % (self.id, d["state"]))
        self.info = d
        self._inputs = None
        self.server._live.pop(self.id, None)
        return d

    def cancel(self):
        """Cancels the job if it is still queued; False once it was admitted."""
        r = self.server.lib.nmk_srv_cancel(self.server.h, self.id)
        if r < 0:
            raise ServingError("unknown job %d" % self.id)
        return r == 0


class Server:
    def __init__(self, max_active=4, hbm_budget=0, device=-1, max_bypass=8):
        self.lib = _lib.load()
        cfg = _lib.SrvConfig()
        self.lib.nmk_srv_default_config(ctypes.byref(cfg))
        cfg.max_active, cfg.hbm_budget, cfg.device, cfg.max_bypass = int(max_active), int(hbm_budget), int(device), \
            int(max_bypass)
        err = _lib.err_buf()
        self.h = self.lib.nmk_srv_create(ctypes.byref(cfg), err, len(err))
        if not self.h:
            raise ServingError(err.value.decode())
        # the server borrows every job's arrays until the job has finished
        self._live = {}

    def submit(self, json_str, lc_vals, mask_vals=None, mask_formula="", precision=32, wl_batch=0, use_graph=False,
               tuning=None, E=None, E_shape=None):
        """Queues a job; returns its :class:`Job`.  Raises ServingError if
        the job is invalid or can never fit the arena."""
        import json as _json
        lc = _lib.as_lc(lc_vals)
        mask = None if mask_vals is None else _lib.as_lc(mask_vals)
        if E is None:
            shape = E_shape
            if shape is None:
                try:
                    shape = fields.output_shape(_json.loads(json_str))
                except Exception:
                    shape = (1,)  # the native settings parser reports what is wrong
            E = np.zeros(shape, dtype=np.complex128)
        if E.dtype != np.complex128 or not E.flags["C_CONTIGUOUS"]:
            raise TypeError("E must be a C-contiguous complex128 array")
        opts = _lib.options(precision=precision, wl_batch=wl_batch, use_graph=use_graph, tuning=tuning)
        err = _lib.err_buf()
        Ef = E.reshape(-1).view(np.float64)
        jid = self.lib.nmk_srv_submit(self.h, json_str.encode(), (mask_formula or "").encode(), _lib.dptr(lc), lc.size,
                                      _lib.dptr(mask), 0 if mask is None else mask.size, _lib.dptr(Ef), E.size,
                                      ctypes.byref(opts), err, len(err))
        if jid < 0:
            raise ServingError(err.value.decode())
        self._live[jid] = (lc, mask, E)
        return Job(self, jid, E, (lc, mask))

    def stats(self):
        c = _lib.SrvCounters()
        self.lib.nmk_srv_stats(self.h, ctypes.byref(c))
        return c.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.nmk_srv_destroy(self.h)  # cancels queued jobs, waits for the running ones
            self.h = None
            self._live = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------- traces
MESHES = [(128, 128, 40), (256, 256, 80), (512, 512, 100), (1024, 1024, 200)]
GENERATORS = ["random-smooth", "droplet", "twisted"]
_SPACING = {128: 0.1, 256: 0.125, 512: 0.1, 1024: 0.05}


def poisson_trace(n_jobs, rate_hz, seed, meshes=MESHES, generators=GENERATORS, max_wl=21, max_q=7):
    """Seeded Poisson arrivals of micrograph jobs: mesh and director generator
    drawn uniformly, 1..max_wl wavelengths over 0.40-0.80 um and the first
    1..max_q incidence wavevectors of the two-ring condenser at NA 0.2."""
    rng = np.random.default_rng(seed)
    ring = fields.ring_wavevectors(0.2, 2)
    t, specs = 0.0, []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / rate_hz)
        mesh = tuple(meshes[rng.integers(len(meshes))])
        gen = generators[rng.integers(len(generators))]
        nwl = int(rng.integers(1, max_wl + 1))
        nq = int(rng.integers(1, max_q + 1))
        wl = [0.6] if nwl == 1 else list(np.linspace(0.40, 0.80, nwl))
        specs.append({"job": i, "arrival_s": t, "mesh": mesh, "generator": gen, "wavelengths": wl,
                      "wavevectors": ring[:nq].tolist()})
    return specs


def director(mesh, generator, cache=None):
    """(lc, mask, mask_formula, spacing) of a trace job, cached per (mesh, generator)."""
    key = (tuple(mesh), generator)
    if cache is not None and key in cache:
        return cache[key]
    Nx, Ny, Nz = mesh
    d = _SPACING.get(Nx, 0.1)
    sp = (d, d, d)
    mask = formula = None
    if generator == "random-smooth":
        lc = fields.random_smooth(Nx, Ny, Nz, sp, seed=5, corr_length=1.0)
    elif generator == "droplet":
        lc, mask, formula = fields.droplet(Nx, Ny, Nz, sp, 0.25 * Nx * d)
    elif generator == "twisted":
        lc = fields.twisted_nematic(Nx, Ny, Nz)
    else:
        raise ValueError("unknown generator " + generator)
    out = (lc, mask, formula, sp)
    if cache is not None:
        cache[key] = out
    return out


def job_settings(spec, spacing):
    return fields.make_settings(spec["mesh"], spacing, spec["wavelengths"], wavevectors=spec["wavevectors"],
                                no="1.5", ne="1.7", nhost="1.51", na_objective=0.4)


def percentiles(values, qs=(50, 99)):
    v = np.asarray(values, dtype=float)
    return {q: float(np.percentile(v, q)) if v.size else float("nan") for q in qs}


def run_trace(server, specs, cache=None, precision=32, wl_batch=0, keep_outputs=False, log=None, use_graph=False):
    """Replays the arrivals of ``specs`` against ``server`` (directors taken
    from / put into ``cache`` before the clock starts) and returns the
    latency percentiles and the aggregate throughput.

    Every time is on the server's clock and measured from the job's scheduled
    arrival (the trace start plus ``arrival_s``), so a late submission counts
    against the job: latency = end - arrival, split into queue (arrival to
    admission), preparation (admission to inputs on the device: mask,
    interface weights, director upload), device (the propagation) and the
    rest (output download)."""
    import json as _json
    cache = {} if cache is None else cache
    prepared = []
    for s in specs:
        lc, mask, formula, sp = director(s["mesh"], s["generator"], cache)
        prepared.append((s, _json.dumps(job_settings(s, sp)), lc, mask, formula))
    jobs = []
    t0 = time.perf_counter()
    srv0 = server.stats()["now_ms"]
    for s, js, lc, mask, formula in prepared:
        delay = s["arrival_s"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        jobs.append((s, server.submit(js, lc, mask_vals=mask, mask_formula=formula or "", precision=precision,
                                      wl_batch=wl_batch, use_graph=use_graph)))
        if log:
            log("submitted job %d %s %s %d wl x %d q" % (s["job"], s["mesh"], s["generator"], len(s["wavelengths"]),
                                                        len(s["wavevectors"])))
    infos = []
    for s, j in jobs:
        info = j.wait()
        if info["state"] != "done":
            raise ServingError("job %d %s: %s" % (s["job"], info["state"], info["error"]))
        info["arrival_ms"] = srv0 + 1e3 * s["arrival_s"]
        infos.append(info)
        if log:
            log("job %d done: latency %.1f ms (queue %.1f, prep %.1f, device %.1f)" % (
                s["job"], info["t_end_ms"] - info["arrival_ms"], info["t_admit_ms"] - info["arrival_ms"],
                info["t_ready_ms"] - info["t_admit_ms"], info["device_ms"]))
    lat = [i["t_end_ms"] - i["arrival_ms"] for i in infos]
    queue = [i["t_admit_ms"] - i["arrival_ms"] for i in infos]
    prep = [i["t_ready_ms"] - i["t_admit_ms"] for i in infos]
    dev = [i["device_ms"] for i in infos]
    lag = [i["t_submit_ms"] - i["arrival_ms"] for i in infos]
    t_first = min(i["arrival_ms"] for i in infos)
    t_last = max(i["t_end_ms"] for i in infos)
    upd = sum(i["updates"] for i in infos)
    p, pq, pp, pd = percentiles(lat), percentiles(queue), percentiles(prep), percentiles(dev)
    st = server.stats()
    out = {"jobs": len(infos), "p50_ms": p[50], "p99_ms": p[99], "mean_ms": float(np.mean(lat)),
           "latency_from": "scheduled arrival",
           "queue_p50_ms": pq[50], "queue_p99_ms": pq[99], "prep_p50_ms": pp[50], "prep_p99_ms": pp[99],
           "device_p50_ms": pd[50], "device_p99_ms": pd[99], "submit_lag_max_ms": float(max(lag)),
           "makespan_s": (t_last - t_first) * 1e-3,
           "updates": upd, "updates_per_s": upd / max(1e-9, (t_last - t_first) * 1e-3),
           "device_busy_ms": float(sum(dev)), "peak_active": st["peak_active"],
           "arena_peak_bytes": st["arena_peak_bytes"], "arena_bytes": st["arena_bytes"]}
    if keep_outputs:
        out["outputs"] = [j.E for _, j in jobs]
    return out
