"""Autotuner of the line-solver launch shapes.

The two cyclic line solvers (x lines and y lines) cut each periodic line into
P chunks and put several lines in one block (rows for x lines, adjacent
columns for y lines).  The best shapes depend on the mesh, the precision and
how many propagations march together, so they are measured on a shallow copy
of the job (same transverse mesh, a few slabs), by coordinate descent: every
x shape that fits the block limits (<= 512 threads, the device's LDS per
block, chunks of at most 16 points; one or both field components per block),
then every y shape (one and two
polarisations per block), x again, the register-lean variants, the eight
XCD work orders, the eight orders with each wavelength's wavevector pairs
grouped, and the slab coefficients on their own stream.  The winner is cached per (device, precision, mesh,
propagation count).
"""
import json
import os
import time

import numpy as np

from . import bpm_backend, fields

_CACHE_ENV = "NEMAKTIS_AMD_TUNING_CACHE"
# LDS of one CDNA4 compute unit; shapes the device refuses are skipped when timed
LDS_MAX = 160 * 1024


def _pow2s(lo, hi):
    v = lo
    while v <= hi:
        yield v
        v *= 2


def _plan(M, chunks):
    cap = max(1, M // 2)
    P = 1
    while P * 2 <= min(cap, chunks):
        P *= 2
    base, rem = M // P, M % P
    need = base + (1 if rem else 0)
    lmax = 4 if need <= 4 else 8 if need <= 8 else 16 if need <= 16 else 0
    return P, lmax


def _lds(M, chunks, lines, c, pols, smap=True):
    """LDS bytes of one line-solver block: the staged right-hand sides and the
    seven reduced-system arrays of each polarisation held by the block
    (x blocks hold both, y blocks one), and the element-to-slot map."""
    P, lmax = _plan(M, chunks)
    return pols * lines * (lmax * P + 7 * P) * c + (M * 4 if smap else 0)


def x_candidates(Mx, precision, both=False):
    """(chunks, rows) of x-line blocks; both: blocks hold both field components"""
    c = 8 if precision == 32 else 16
    k = 4 if both else 2
    out = []
    for chunks in _pow2s(16, 512):
        P, lmax = _plan(Mx, chunks)
        if lmax == 0 or lmax > 16 or P != chunks:
            continue
        for rows in (1, 2, 4, 8):
            if k * P * rows <= 512 and _lds(Mx, chunks, rows, c, k, True) <= LDS_MAX:
                out.append((chunks, rows))
    return out


def y_candidates(My, precision, pair=False):
    c = 8 if precision == 32 else 16
    k = 2 if pair else 1
    out = []
    for chunks in _pow2s(16, 512):
        P, lmax = _plan(My, chunks)
        if lmax == 0 or lmax > 16 or P != chunks:
            continue
        for cols in (1, 2, 4, 8, 16, 32):
            if k * P * cols <= 512 and _lds(My, chunks, cols, c, k, True) <= LDS_MAX:
                out.append((chunks, cols))
    return out


def _cache_path():
    p = os.environ.get(_CACHE_ENV)
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "nemaktis_amd", "tuning.json")


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except Exception:
        return {}


def _save_cache(c):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
    except OSError:
        pass


def cache_key(precision, Nx, Ny, nprops, use_graph=True):
    return "gfx950|solver-v13|%d|%dx%d|%d|%s" % (precision, Nx, Ny, nprops, "graph" if use_graph else "eager")


def _time(ctx, tuning, reps):
    """Median march time of the resident job with `tuning` (inf when the
    device refuses the shape).  One context serves every candidate: buffers
    that a captured graph has used are never freed between candidates."""
    try:
        ctx.set_tuning(tuning)
    except bpm_backend.BpmBackendError:
        return float("inf")  # shape beyond the device limits
    ctx.propagate()
    t = []
    for _ in range(reps):
        t.append(ctx.propagate()["march_ms"])
    return float(np.median(t))


def tune(settings, lc_vals, precision=32, wl_batch=0, depth=9, reps=3, use_cache=True, log=None, use_graph=True):
    """Returns the best tuning dict for this job's transverse mesh, measured
    on a `depth`-plane copy of it (graph replay or eager launches, as the job
    will run), plus the timing table."""
    ic = settings["Physics settings"]["Initial conditions"]
    Nx, Ny, Nz = ic["Mesh dimensions"]
    co = settings["Physics settings"]["Coefficients"]
    nwl = len(co["Wavelengths"])
    nprops = (nwl if wl_batch <= 0 else min(wl_batch, nwl)) * (len(co["Wavevectors"]) // 2) * 2
    key = cache_key(precision, Nx, Ny, nprops, use_graph)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        return cache[key]["tuning"], cache[key].get("table", [])
    depth = max(2, min(depth, Nz))
    s = json.loads(json.dumps(settings))
    s["Physics settings"]["Initial conditions"]["Mesh dimensions"] = [Nx, Ny, depth]
    lc = np.ascontiguousarray(np.asarray(lc_vals).reshape(Nz, Ny, Nx, -1)[:depth])
    js = json.dumps(s)
    n_E = int(np.prod(fields.output_shape(s)))
    Mx, My = (Nx - 1 if Nx > 3 else Nx), (Ny - 1 if Ny > 3 else Ny)
    xs, ys = x_candidates(Mx, precision), y_candidates(My, precision)
    if not xs or not ys:
        return None, []
    c = 8 if precision == 32 else 16
    x0 = next((x for x in xs if _lds(Mx, x[0], x[1], c, 2, True) <= 64 * 1024), xs[0])
    y0 = next((y for y in ys if _lds(My, y[0], y[1], c, 1, True) <= 64 * 1024), ys[0])
    best = {"x_chunks": x0[0], "x_rows": x0[1], "y_chunks": y0[0], "y_cols": y0[1], "coef_serial": 1}
    ctx = bpm_backend.Context(js, lc, n_E, precision=precision, wl_batch=wl_batch, use_graph=use_graph, tuning=best)
    try:
        best, best_ms, table, t0 = _search(ctx, best, xs, ys, Mx, My, precision, reps)
    finally:
        ctx.close()
    if log:
        log("autotune %s: %d candidates in %.1f s, best %s (%.3f ms for %d slabs)"
            % (key, len(table), time.time() - t0, best, best_ms, depth - 1))
        for c, ms in sorted(table, key=lambda cm: cm[1])[:8]:
            log("   %.3f ms  %s" % (ms, c))
    if use_cache:
        cache[key] = {"tuning": best, "table": [[c, m] for c, m in table]}
        _save_cache(cache)
    return best, table


def _search(ctx, best, xs, ys, Mx, My, precision, reps):
    """Coordinate descent over the launch shapes on one resident job."""
    table = []
    t0 = time.time()
    best_ms = None
    # x_both bit 0: first launches hold both components, bit 1: Neumann launches
    xs2 = x_candidates(Mx, precision, both=True)
    for both, shapes in ((0, xs), (1, xs2), (3, xs2)):
        for (xc, xr) in shapes:
            cand = dict(best, x_chunks=xc, x_rows=xr, x_both=both)
            ms = _time(ctx, cand, reps)
            table.append((cand, ms))
            if best_ms is None or ms < best_ms:
                best, best_ms = cand, ms
    for (yc, yw) in ys:
        cand = dict(best, y_chunks=yc, y_cols=yw)
        ms = _time(ctx, cand, reps)
        table.append((cand, ms))
        if ms < best_ms:
            best, best_ms = cand, ms
    # y blocks holding both polarisations (one coefficient gather for the two),
    # or both field components (whose right-hand sides have unequal tap counts)
    for pair in (1, 2):
        for (yc, yw) in y_candidates(My, precision, pair=True):
            if yw < 4:
                continue
            cand = dict(best, y_chunks=yc, y_cols=yw, y_pair=pair)
            ms = _time(ctx, cand, reps)
            table.append((cand, ms))
            if ms < best_ms:
                best, best_ms = cand, ms
    # second pass over the x shapes with the y shape found (coordinate descent)
    for both, shapes in ((0, xs), (1, xs2), (2, xs2), (3, xs2)):
        for (xc, xr) in shapes:
            cand = dict(best, x_chunks=xc, x_rows=xr, x_both=both)
            if cand == best:
                continue
            ms = _time(ctx, cand, reps)
            table.append((cand, ms))
            if ms < best_ms:
                best, best_ms = cand, ms
    # register-lean variants (more waves per SIMD) with the shapes found so far
    # and with shorter chunks
    for (xc, yc) in [(best["x_chunks"], best["y_chunks"]), (2 * best["x_chunks"], 2 * best["y_chunks"])]:
        cand = dict(best, x_chunks=xc, y_chunks=yc, lean=1)
        xok = x_candidates(Mx, precision, both=bool(cand.get("x_both")))
        if (xc, cand["x_rows"]) not in xok or (yc, cand["y_cols"]) not in y_candidates(My, precision,
                                                                                      pair=bool(cand.get("y_pair"))):
            continue
        ms = _time(ctx, cand, reps)
        table.append((cand, ms))
        if ms < best_ms:
            best, best_ms = cand, ms
    # joint pass over the x block contents (components, rows) and the XCD work
    # orders (which items share an XCD's L2: components, neighbours,
    # wavevectors), which interact through L2 locality and load balance
    for both, shapes in ((0, xs), (1, xs2), (2, xs2), (3, xs2)):
        for (xc, xr) in shapes:
            if xc != best["x_chunks"]:
                continue
            for order in range(8):
                cand = dict(best, x_rows=xr, x_both=both, order=order)
                if cand == best:
                    continue
                ms = _time(ctx, cand, reps)
                table.append((cand, ms))
                if ms < best_ms:
                    best, best_ms = cand, ms
    # the wavevector pairs of each wavelength grouped (orders 8-15) with the
    # shapes found
    for order in range(8, 16):
        cand = dict(best, order=order)
        ms = _time(ctx, cand, reps)
        table.append((cand, ms))
        if ms < best_ms:
            best, best_ms = cand, ms
    # slab coefficients on a second stream, overlapping the march (at the cost
    # of two stream joins per slab)
    cand = dict(best, coef_serial=0)
    ms = _time(ctx, cand, reps)
    table.append((cand, ms))
    if ms < best_ms:
        best, best_ms = cand, ms
    return best, best_ms, table, t0
