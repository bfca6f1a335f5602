"""GPU parity of the HIP path against the FP64 CPU oracle, through the C ABI.

Tolerances (relative L2 per output (wavelength, wavevector, polarisation)):
  FP64 path: 1e-10      FP32 path: 1e-4
Batched (all wavelengths at once) and hipGraph-replayed runs must equal the
plain run bit for bit.
"""
import json
import math

import numpy as np
import pytest

import oracle
from nemaktis_amd import bpm_backend, fields

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-4


def _gpu(s, lc, mask=None, formula="", E_in=None, **kw):
    print("[gpu case]", s["Physics settings"]["Initial conditions"]["Mesh dimensions"], kw, flush=True)
    return bpm_backend.propagate(json.dumps(s), lc, mask_vals=mask, mask_formula=formula, E_in=E_in, **kw)


def _rel(a, b):
    """max over outputs of the relative L2 error of a against b"""
    a = a.reshape(a.shape[0], a.shape[1], 2, -1)
    b = b.reshape(b.shape[0], b.shape[1], 2, -1)
    num = np.linalg.norm(a - b, axis=-1)
    den = np.linalg.norm(b, axis=-1)
    return float((num / den).max())


def _formula_weights(s, formula_fn, mask):
    """Interface weights by bisection (UtilitiesLib/VectorField.cpp:64-113)."""
    Nx, Ny, Nz = s["Physics settings"]["Initial conditions"]["Mesh dimensions"]
    dx, dy, dz = s["Physics settings"]["Initial conditions"]["Mesh spacings"]
    w = np.full((Nz - 1, Ny, Nx, 2), 0.5)
    m = (mask >= 0)
    failed = False
    for iy in range(Ny):
        for ix in range(Nx):
            x, y = (ix - 0.5 * (Nx - 1)) * dx, (iy - 0.5 * (Ny - 1)) * dy
            f = lambda z: formula_fn(x, y, z)
            for iz in range(Nz - 1):
                zi = (iz - 0.5 * (Nz - 1)) * dz
                zm, zp = zi, (iz + 1 - 0.5 * (Nz - 1)) * dz
                if f(zm) * f(zp) <= 0:
                    if m[iz, iy, ix] == m[iz + 1, iy, ix]:
                        failed = True
                    z = zm
                    while abs(zp - zm) >= 1e-3 * dz:
                        z = zm
                        fm = f(z)
                        z = 0.5 * (zp + zm)
                        fz = f(z)
                        if fz == 0:
                            break
                        if fz * fm < 0:
                            zp = z
                        else:
                            zm = z
                    w[iz, iy, ix, 0] = (z - zi) / dz
                    w[iz, iy, ix, 1] = 1 - w[iz, iy, ix, 0]
    return None if failed else w


def _base(dims, sp=(0.1, 0.1, 0.1), **kw):
    return fields.make_settings(dims, sp, kw.pop("wl", [0.6]), **kw)


def test_twisted_nematic_fp64():
    s = _base((16, 16, 6), no="1.5", ne="1.7", nhost="1.5")
    lc = fields.twisted_nematic(16, 16, 6)
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


def _random_case(dims=(20, 18, 6), seed=3):
    s = _base(dims, wl=[0.55, 0.62], wavevectors=[(0.0, 0.0), (0.15, -0.1)], no="1.5",
              ne="1.7+0.01/lambda^2", ne_imag="0.01", nin="1.2", nout="1.1", na_objective=0.7)
    lc = fields.random_smooth(dims[0], dims[1], dims[2], (0.1, 0.1, 0.1), seed=seed, corr_length=0.3)
    return s, lc


def test_random_director_walkoff_shift_fp64():
    s, lc = _random_case()
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


@pytest.mark.parametrize("nw", [0, 1, 4])
def test_neumann_depth_fp64(nw):
    s, lc = _random_case((14, 12, 4))
    s["Algorithm settings"]["Beam propagation"]["N Woodbury steps"] = nw
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


def test_droplet_mask_gaussian_beam_fp64():
    dims, sp = (24, 22, 8), (0.25, 0.25, 0.25)
    lc, mask, _ = fields.droplet(*dims, sp, 2.2)
    s = _base(dims, sp, wavevectors=[(0.0, 0.0), (0.1, 0.05)], no="1.5", ne="1.7", nhost="1.51",
              beam_profile="GaussianBeam(1.5)", na_objective=0.5)
    err = _rel(_gpu(s, lc, mask=mask), oracle.propagate(s, lc, mask=mask))
    assert err < TOL64, err


def test_droplet_mask_formula_weights_fp64():
    dims, sp = (20, 20, 9), (0.25, 0.25, 0.25)
    R = 1.9
    lc, mask, formula = fields.droplet(*dims, sp, R)
    s = _base(dims, sp, no="1.5", ne="1.7", nhost="1.51", na_objective=0.5)
    w = _formula_weights(s, lambda x, y, z: R * R - x * x - y * y - z * z, mask)
    assert w is not None
    want = oracle.propagate(s, lc, mask=mask, weights=w)
    err = _rel(_gpu(s, lc, mask=mask, formula=formula), want)
    assert err < TOL64, err
    # the weights matter: without them the result differs
    assert _rel(oracle.propagate(s, lc, mask=mask), want) > 1e-6


def test_qtensor_field_fp64():
    dims = (16, 14, 5)
    n = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=9, corr_length=0.3)
    Q = 0.6 * (np.einsum("...i,...j->...ij", n, n) - np.eye(3) / 3)
    lc = np.stack([Q[..., 0, 0], Q[..., 1, 1], Q[..., 2, 2], Q[..., 0, 1], Q[..., 0, 2], Q[..., 1, 2]], -1)
    s = _base(dims, no="1.5", ne="1.7", field_type="Q-tensor", wavevectors=[(0.05, 0.1)])
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


def test_iso_layers_focus_and_caller_fields_fp64():
    s, lc = _random_case((16, 16, 5), seed=5)
    s["Postprocessor settings"]["Screen output"].update({
        "Isotropic layer thicknesses": [1.0, 0.5], "Isotropic layer refractive indices": [1.45, 1.3],
        "Focalisation z-shift": 0.7})
    s["Physics settings"]["Initial conditions"]["Beam profile"] = "None"
    rng = np.random.default_rng(1)
    E_in = rng.standard_normal(fields.output_shape(s)) + 1j * rng.standard_normal(fields.output_shape(s))
    err = _rel(_gpu(s, lc, E_in=E_in), oracle.propagate(s, lc, E=E_in))
    assert err < TOL64, err


def test_dispersive_nout_with_iso_layers_and_focus_fp64():
    """A dispersive exit index with isotropic layers and a focus shift over
    three wavelengths: the focus correction uses nout at 0.6 um
    (PostprocessorLib/ScreenOutput.cpp:34-38) while each wavelength's filter
    uses its own nout, as the oracle restates; FP64 and FP32 against it."""
    dims = (18, 14, 5)
    s = _base(dims, wl=[0.45, 0.6, 0.75], wavevectors=[(0.0, 0.0), (0.06, -0.03)], no="1.5",
              ne="1.7+0.01/lambda^2", nout="1.33+0.004/lambda^2", na_objective=0.6)
    s["Postprocessor settings"]["Screen output"].update({
        "Isotropic layer thicknesses": [0.8, 0.3], "Isotropic layer refractive indices": [1.45, 1.38],
        "Focalisation z-shift": 0.9})
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=73, corr_length=0.3)
    want = oracle.propagate(s, lc)
    assert _rel(_gpu(s, lc), want) < TOL64
    assert _rel(_gpu(s, lc, precision=32, wl_batch=0), want) < TOL32


@pytest.mark.parametrize("dims", [(15, 12, 5), (3, 20, 4), (18, 1, 4), (21, 10, 3), (12, 10, 2), (2, 2, 3)])
def test_odd_and_degenerate_meshes_fp64(dims):
    s = _base(dims, wavevectors=[(0.08, -0.12)], no="1.5", ne="1.7", na_objective=0.6)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=11, corr_length=0.3)
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


def test_full_twisted_nematic_case_fp64():
    s, lc, _, _ = fields.workload("tn64")
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


def test_fp32_path_error_bound():
    s, lc = _random_case((48, 40, 12), seed=7)
    want = oracle.propagate(s, lc)
    err = _rel(_gpu(s, lc, precision=32), want)
    print("fp32 rel L2", err)
    assert err < TOL32


def test_batched_and_graph_equal_plain_bitwise():
    s, lc = _random_case((18, 16, 5), seed=13)
    plain = _gpu(s, lc)
    batched = _gpu(s, lc, wl_batch=0)
    assert np.array_equal(plain, batched)
    js = json.dumps(s)
    n_E = int(np.prod(fields.output_shape(s)))
    with bpm_backend.Context(js, lc, n_E, wl_batch=0, use_graph=True) as ctx:
        for _ in range(3):
            ctx.propagate()
            out = ctx.download(np.empty(fields.output_shape(s), np.complex128))
            assert np.array_equal(out, plain)
    with bpm_backend.Context(js, lc, n_E, precision=32, wl_batch=0, use_graph=True) as ctx:
        outs = []
        for _ in range(3):
            ctx.propagate()
            outs.append(ctx.download(np.empty(fields.output_shape(s), np.complex128)))
        assert all(np.array_equal(o, outs[0]) for o in outs)
        err = _rel(outs[0], plain)
    assert err < TOL32, err


def test_random512_reduced_depth_fp32_vs_fp64():
    """512x512 slabs (the benchmark's transverse size, 8 slabs deep): FP32 vs FP64."""
    s, _, _, _ = fields.workload("tn64")
    s = _base((512, 512, 9), no="1.52", ne="1.74", wl=[0.55], na_objective=0.4)
    lc = fields.random_smooth(512, 512, 9, (0.1, 0.1, 0.1), seed=2, corr_length=1.0)
    e64 = _gpu(s, lc)
    e32 = _gpu(s, lc, precision=32)
    assert np.isfinite(e64).all()
    err = _rel(e32, e64)
    assert err < TOL32, err


# launch shapes of the autotuner's candidate forms on 69 x 65-point lines
SOLVER_SHAPES = [
    {"x_chunks": 8, "x_rows": 4, "y_chunks": 8, "y_cols": 4},     # chunks of 8-9 points (16-point solver)
    {"x_chunks": 32, "x_rows": 1, "y_chunks": 32, "y_cols": 1},   # chunks of 2-3 points
    {"x_chunks": 16, "x_rows": 2, "y_chunks": 16, "y_cols": 8},   # chunks of 4-5 points
    {"x_chunks": 16, "x_rows": 2, "y_chunks": 16, "y_cols": 8, "y_pair": 1, "lean": 1},  # paired y, lean
    {"x_chunks": 16, "x_rows": 1, "y_chunks": 16, "y_cols": 4, "x_both": 3, "y_pair": 2},  # both components
    {"x_chunks": 8, "x_rows": 2, "y_chunks": 8, "y_cols": 2, "x_both": 2, "y_pair": 2, "lean": 1},
    {"x_chunks": 32, "x_rows": 4, "y_chunks": 16, "y_cols": 8, "x_both": 3},  # four rows of both components
    {"x_chunks": 16, "x_rows": 1, "y_chunks": 16, "y_cols": 8, "x_both": 1},  # both components, first launches only
    {"x_chunks": 16, "x_rows": 2, "y_chunks": 16, "y_cols": 8, "coef_serial": 1},  # coefficients on the march stream
]


@pytest.mark.parametrize("tuning", SOLVER_SHAPES)
def test_line_solver_shapes_fp64(tuning):
    """Every autotuner candidate shape must give the oracle's answer."""
    s, lc = _random_case((70, 66, 5), seed=17)
    want = oracle.propagate(s, lc)
    err = _rel(_gpu(s, lc, tuning=tuning), want)
    assert err < TOL64, err


def test_builtin_shapes_mid_mesh_batched():
    """Built-in launch shapes on a mesh with 299- and 259-point lines and 12
    propagations (32 chunks per line, 16-column y strips, pairs-innermost
    work order), FP32 graph replay and FP64, against the oracle."""
    s = _base((300, 260, 4), wl=[0.5, 0.6, 0.7], wavevectors=[(0.0, 0.0), (0.1, 0.05)], no="1.5", ne="1.7",
              nhost="1.5", na_objective=0.4)
    lc = fields.random_smooth(300, 260, 4, (0.1, 0.1, 0.1), seed=31, corr_length=0.3)
    want = oracle.propagate(s, lc)
    with bpm_backend.Context(json.dumps(s), lc, int(np.prod(fields.output_shape(s))), precision=32, wl_batch=0,
                             use_graph=True) as ctx:
        ctx.propagate()
        got32 = ctx.download(np.empty(fields.output_shape(s), np.complex128))
    err32 = _rel(got32, want)
    assert err32 < TOL32, err32
    err64 = _rel(_gpu(s, lc, precision=64, wl_batch=0), want)
    assert err64 < TOL64, err64


def test_both_component_blocks_with_lds_reduced_systems():
    """x lines of 299 points in 128 chunks (reduced systems in LDS) with both
    components per block, against 64 chunks (lane permutes), FP64."""
    s, lc = _random_case((300, 40, 4), seed=23)
    a = _gpu(s, lc, tuning={"x_chunks": 128, "x_rows": 1, "y_chunks": 16, "y_cols": 4, "x_both": 3})
    b = _gpu(s, lc, tuning={"x_chunks": 64, "x_rows": 2, "y_chunks": 16, "y_cols": 4})
    assert _rel(a, b) < 1e-12


def test_retuned_context_matches_fresh_context():
    """nmk_bpm_set_tuning on a resident job (graph replay) gives what a
    context created with that tuning gives, and a refused shape leaves the
    context usable."""
    s, lc = _random_case((40, 36, 5), seed=29)
    js = json.dumps(s)
    shape = fields.output_shape(s)
    n_E = int(np.prod(shape))
    t1 = {"x_chunks": 16, "x_rows": 2, "y_chunks": 16, "y_cols": 8}
    t2 = {"x_chunks": 8, "x_rows": 1, "y_chunks": 8, "y_cols": 4, "x_both": 3, "y_pair": 2, "order": 5,
          "coef_serial": 1}
    with bpm_backend.Context(js, lc, n_E, precision=32, wl_batch=0, use_graph=True, tuning=t1) as ctx:
        ctx.propagate()
        ctx.set_tuning(t2)
        ctx.propagate()
        got = ctx.download(np.empty(shape, np.complex128))
        with pytest.raises(bpm_backend.BpmBackendError, match="block limits"):
            ctx.set_tuning({"x_chunks": 16, "x_rows": 64, "y_chunks": 16, "y_cols": 8})
        ctx.propagate()
        again = ctx.download(np.empty(shape, np.complex128))
    with bpm_backend.Context(js, lc, n_E, precision=32, wl_batch=0, use_graph=True, tuning=t2) as ctx:
        ctx.propagate()
        want = ctx.download(np.empty(shape, np.complex128))
    assert np.array_equal(got, want) and np.array_equal(again, want)


def test_graph_contexts_return_their_memory():
    """Graph-mode contexts created, replayed and destroyed one after another
    give their device memory back: after one warm-up context (the runtime's
    own one-time graph state), the free memory the runtime reports does not
    drift over three more."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    s = _base((128, 128, 8), wl=[0.5, 0.6], no="1.5", ne="1.7", na_objective=0.4)
    lc = fields.random_smooth(128, 128, 8, (0.1, 0.1, 0.1), seed=3, corr_length=0.3)
    js, n_E = json.dumps(s), int(np.prod(fields.output_shape(s)))
    free = []
    for i in range(4):
        with bpm_backend.Context(js, lc, n_E, precision=32, wl_batch=0, use_graph=True) as ctx:
            ctx.propagate()
            ctx.propagate()
            held = free_bytes()
        free.append(free_bytes())
        assert free[-1] - held > 16 << 20      # the context's buffers came back
    print("free memory after each context (MB):", [f >> 20 for f in free])
    assert max(free[1:]) - min(free[1:]) <= 4 << 20


def test_xcd_work_orders_are_bitwise_equal():
    """The XCD work orders only move blocks between XCDs: identical output
    (two wavelengths x two wavevectors, so that every index varies; orders 8-15
    group the two pairs of each wavelength)."""
    s, lc = _random_case((70, 66, 5), seed=19)
    base = {"x_chunks": 16, "x_rows": 2, "y_chunks": 16, "y_cols": 8}
    outs = [_gpu(s, lc, precision=32, wl_batch=0, tuning=dict(base, order=o)) for o in range(16)]
    outs += [_gpu(s, lc, precision=32, wl_batch=0, tuning=dict(base, order=o, y_pair=1)) for o in (0, 6, 8, 14)]
    outs += [_gpu(s, lc, precision=32, wl_batch=0, tuning=dict(base, order=o, x_both=3, x_rows=1, y_pair=2,
                                                                 y_cols=4)) for o in (1, 4)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    want = oracle.propagate(s, lc)
    assert _rel(outs[0], want) < TOL32


def test_drop_in_c_symbols_called_directly():
    """nmk_bpm_run_backend_without_mask / _with_mask through ctypes, as a C
    caller binds them (PythonBackend/bpm_backend.h:4-13), against the oracle."""
    import ctypes
    from nemaktis_amd import _lib
    lib = _lib.load()
    s, lc = _random_case((20, 18, 6), seed=41)
    js = json.dumps(s).encode()
    lcf = np.ascontiguousarray(lc.reshape(-1))
    E = np.zeros(fields.output_shape(s), np.complex128)
    err = _lib.err_buf()
    print("[gpu case] nmk_bpm_run_backend_without_mask", flush=True)
    rc = lib.nmk_bpm_run_backend_without_mask(js, _lib.dptr(lcf), lcf.size, _lib.dptr(E.reshape(-1).view(np.float64)),
                                              E.size, err, len(err))
    assert rc == 0 and err.value == b"", err.value
    assert _rel(E, oracle.propagate(s, lc)) < TOL64
    dims, sp = (20, 20, 9), (0.25, 0.25, 0.25)
    R = 1.9
    lcd, mask, formula = fields.droplet(*dims, sp, R)
    sd = _base(dims, sp, no="1.5", ne="1.7", nhost="1.51", na_objective=0.5)
    w = _formula_weights(sd, lambda x, y, z: R * R - x * x - y * y - z * z, mask)
    E2 = np.zeros(fields.output_shape(sd), np.complex128)
    m = np.ascontiguousarray(mask.reshape(-1))
    lcdf = np.ascontiguousarray(lcd.reshape(-1))
    print("[gpu case] nmk_bpm_run_backend_with_mask", flush=True)
    rc = lib.nmk_bpm_run_backend_with_mask(json.dumps(sd).encode(), formula.encode(), _lib.dptr(lcdf), lcdf.size,
                                           _lib.dptr(m), m.size, _lib.dptr(E2.reshape(-1).view(np.float64)), E2.size,
                                           err, len(err))
    assert rc == 0 and err.value == b"", err.value
    assert _rel(E2, oracle.propagate(sd, lcd, mask=mask, weights=w)) < TOL64


def test_mask_formula_mismatch_notice_and_fallback(capsys):
    """A mask formula that disagrees with the mask: the reference's notice on
    stdout and the default weights (UtilitiesLib/VectorField.cpp:107-111)."""
    dims, sp = (16, 14, 6), (0.25, 0.25, 0.25)
    lc, mask, _ = fields.droplet(*dims, sp, 1.2)
    s = _base(dims, sp, no="1.5", ne="1.7", nhost="1.51", na_objective=0.5)
    E = np.zeros(fields.output_shape(s), np.complex128)
    bpm_backend.run_backend_with_mask(json.dumps(s), "0.1-z", lc, mask, E, strict=True)
    out = capsys.readouterr().out
    assert "Mask values do not corresponds to the given formula, will ignore interpolation correction" in out
    assert _rel(E, oracle.propagate(s, lc, mask=mask)) < TOL64


def test_many_iso_layers_fp64():
    """More than 16 isotropic layers between the sample and the objective
    (PostprocessorLib/ScreenOutput.cpp:240-299 loops over any number)."""
    s, lc = _random_case((16, 14, 4), seed=43)
    n = 20
    s["Postprocessor settings"]["Screen output"].update({
        "Isotropic layer thicknesses": [0.05 + 0.01 * k for k in range(n)],
        "Isotropic layer refractive indices": [1.3 + 0.01 * k for k in range(n)],
        "Focalisation z-shift": 0.2})
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err


@pytest.mark.parametrize("dims", [(17, 16, 4), (34, 19, 4), (16, 97, 3)])
def test_mesh_sizes_with_large_prime_factors_fp64(dims):
    """Screen FFT of lengths with prime factors 17, 19 and 97 (direct DFT on the
    device; FFTW plans any length, BeamPropagationLib/OpticalFieldCollection.cpp:59-71)
    against the oracle (Bluestein)."""
    s = _base(dims, wavevectors=[(0.0, 0.0), (0.07, -0.04)], no="1.5", ne="1.7", na_objective=0.6)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=47, corr_length=0.3)
    err = _rel(_gpu(s, lc), oracle.propagate(s, lc))
    assert err < TOL64, err
    err32 = _rel(_gpu(s, lc, precision=32), oracle.propagate(s, lc))
    assert err32 < TOL32, err32


@pytest.mark.parametrize("dims", [(3000, 6, 3), (8, 2600, 3)])
def test_lines_longer_than_2048_points_fp32(dims):
    """Lines of 2999 and 2599 unknowns (256 chunks per line, reduced systems in
    LDS) and screen transforms of 3000 and 2600 points (direct DFT), FP32."""
    s = _base(dims, no="1.5", ne="1.7", na_objective=0.5)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=53, corr_length=0.3)
    err = _rel(_gpu(s, lc, precision=32), oracle.propagate(s, lc))
    assert err < TOL32, err


@pytest.mark.parametrize("dims,precision", [((4097, 5, 3), 32), ((5, 4097, 3), 32), ((2049, 5, 3), 64),
                                            ((5, 2049, 3), 64)])
def test_maximum_line_lengths(dims, precision):
    """The longest lines each path takes: 4097 points (4096 unknowns, 256
    chunks of 16) in FP32 and 2049 (256 chunks of 8) in FP64, along x and
    along y, against the oracle; one point more is refused when the settings
    are read."""
    s = _base(dims, no="1.5", ne="1.7", na_objective=0.5)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=59, corr_length=0.3)
    print("[gpu case] maximum line length", dims, precision, flush=True)
    err = _rel(_gpu(s, lc, precision=precision), oracle.propagate(s, lc))
    assert err < (TOL32 if precision == 32 else TOL64), err
    big = tuple(d + 1 if d > 5 else d for d in dims)
    s2 = _base(big, no="1.5", ne="1.7", na_objective=0.5)
    lc2 = fields.random_smooth(*big, (0.1, 0.1, 0.1), seed=59, corr_length=0.3)
    with pytest.raises(bpm_backend.BpmBackendError, match=str(max(dims))):
        _gpu(s2, lc2, precision=precision)


def test_single_plane_mesh():
    """A mesh of one plane (Nz = 1; the reference's loop over Nz-1 slabs runs
    zero times): entrance and exit interfaces and the screen filter only, each
    point of the duplicate column and row with its own permittivity; FP64 and
    FP32 against the oracle, with and without a mask, caller-supplied fields,
    and a graph context (nothing to capture) equal to the one-shot run."""
    dims, sp = (20, 14, 1), (0.1, 0.1, 0.1)
    s = _base(dims, sp, no="1.5", ne="1.7", nhost="1.45", nin="1.2", nout="1.33", na_objective=0.5,
              wavevectors=[(0.0, 0.0), (0.05, 0.02)])
    lc = fields.random_smooth(*dims, sp, seed=61, corr_length=0.3)
    want = oracle.propagate(s, lc)
    assert _rel(_gpu(s, lc), want) < TOL64
    assert _rel(_gpu(s, lc, precision=32), want) < TOL32
    mask = np.where(np.arange(dims[0])[None, None, :] < 10, 1.0, -1.0) * np.ones(dims[::-1])
    assert _rel(_gpu(s, lc, mask=mask), oracle.propagate(s, lc, mask=mask)) < TOL64
    sn = _base(dims, sp, no="1.5", ne="1.7", nhost="1.45", nout="1.33", na_objective=0.5, beam_profile="None")
    rng = np.random.default_rng(67)
    shape = fields.output_shape(sn)
    Ein = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert _rel(_gpu(sn, lc, E_in=Ein), oracle.propagate(sn, lc, E=Ein)) < TOL64
    one = _gpu(s, lc, precision=32, wl_batch=0)
    shape = fields.output_shape(s)
    with bpm_backend.Context(json.dumps(s), lc, int(np.prod(shape)), precision=32, wl_batch=0,
                             use_graph=True) as ctx:
        ctx.propagate()
        got = ctx.download(np.empty(shape, np.complex128))
    assert np.array_equal(got, one)


def test_staged_and_direct_host_copies_agree(tmp_path):
    """Host copies of 16 MB and more go through pinned chunks filled by host
    threads; a child process with NMK_NO_STAGING=1 copies from the caller's
    memory directly.  Same output bit for bit (a 126 MB director in, 8 MB
    out; a 617 MB output out on the droplet sweep is covered by the
    benchmark-workload tests)."""
    import os
    import subprocess
    import sys
    dims = (256, 256, 80)
    s = _base(dims, no="1.5", ne="1.7", na_objective=0.4)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=71, corr_length=0.5)
    np.save(tmp_path / "lc.npy", lc)
    (tmp_path / "s.json").write_text(json.dumps(s))
    staged = _gpu(s, lc, precision=32, wl_batch=0)
    code = ("import json, sys, numpy as np; sys.path.insert(0, %r); from nemaktis_amd import bpm_backend; "
            "d = %r; s = open(d + '/s.json').read(); lc = np.load(d + '/lc.npy'); "
            "np.save(d + '/E.npy', bpm_backend.propagate(s, lc, precision=32, wl_batch=0))"
            % (os.path.dirname(os.path.dirname(os.path.abspath(__file__))), str(tmp_path)))
    print("[gpu case] direct host copies in a child process", flush=True)
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, NMK_NO_STAGING="1"), timeout=240,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert np.array_equal(np.load(tmp_path / "E.npy"), staged)


def test_batched_slab_coefficients_with_a_partial_last_batch():
    """A small mesh computes the coefficients of several slabs per launch (a ring of twice the batch): 22 slabs
    of a 40 x 36 mesh come in batches of 8, 8 and 6, with shifted wavevectors (the shift boundary reads two
    slabs' phase planes) and the Neumann depth 2: one-shot FP64 (one wavelength at a time) and FP32 (all
    together), and graph-replaying contexts propagated twice."""
    s, lc = _random_case((40, 36, 23), seed=29)
    s["Physics settings"]["Coefficients"]["Wavelengths"] = [0.55, 0.65]
    s["Physics settings"]["Coefficients"]["Wavevectors"] = [0.0, 0.0, 0.12, -0.05]
    want = oracle.propagate(s, lc)
    assert _rel(_gpu(s, lc, precision=64), want) < TOL64
    assert _rel(_gpu(s, lc, precision=32, wl_batch=0), want) < TOL32
    n = int(np.prod(fields.output_shape(s)))
    for prec, tol in ((64, TOL64), (32, TOL32)):
        print("[gpu case] batched slab coefficients, graph context, precision %d" % prec, flush=True)
        with bpm_backend.Context(json.dumps(s), lc, n, precision=prec, wl_batch=0, use_graph=True) as ctx:
            outs = []
            for _ in range(2):
                ctx.propagate()
                outs.append(ctx.download(np.empty(fields.output_shape(s), np.complex128)))
        assert np.array_equal(outs[0], outs[1]) and _rel(outs[0], want) < tol
