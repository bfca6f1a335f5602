"""CPU tests of the host side of the drop-in boundary: the C-ABI library loads
and exports what include/*.h declares, the settings schema and error
behaviour mirror the reference (BeamPropagationSolver/src/SettingsLib/*,
PythonBackend/bpm_backend.cpp), and the index-expression evaluator follows
muParser syntax.  No compute call reaches a GPU here."""
import ctypes
import glob
import io
import json
import math
import os
import re
import sys

import numpy as np
import pytest

from nemaktis_amd import _lib, bpm_backend, fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(nmk_\w+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = _declared_symbols()
    assert "nmk_bpm_run_backend_without_mask" in decl and "nmk_bpm_run_backend_with_mask" in decl
    for name in decl:
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTED) <= set(decl)
    assert b"gfx950" in lib.nmk_bpm_version()


def _eval(expr, lam=0.6):
    v = ctypes.c_double()
    err = _lib.err_buf()
    rc = _lib.load().nmk_bpm_eval_expression(expr.encode(), lam, ctypes.byref(v), err, len(err))
    if rc:
        raise ValueError(err.value.decode())
    return v.value


@pytest.mark.parametrize("expr,lam,want", [
    ("1.5", 0.6, 1.5),
    ("1.5+0.01/lambda^2", 0.5, 1.5 + 0.01 / 0.25),
    ("-2^2", 0.6, -4.0),
    ("2^3^2", 0.6, 2.0 ** 9),
    ("sqrt(2.25)*(1+0*lambda)", 0.6, 1.5),
    ("lambda<0.5 ? 1.6 : 1.5", 0.4, 1.6),
    ("lambda<0.5 ? 1.6 : 1.5", 0.7, 1.5),
    ("1.7+0.0158/lambda^2+0.00167/lambda^4", 0.55, 1.7 + 0.0158 / 0.55 ** 2 + 0.00167 / 0.55 ** 4),
    ("max(1.5, lambda, 2)", 3.0, 3.0),
    ("_pi", 0.6, math.pi),
    ("exp(ln(2))", 0.6, 2.0),
    ("1e-3*2.5E2", 0.6, 0.25),
    (" 3 * ( 2 - lambda ) ", 0.5, 4.5),
])
def test_expression_evaluator(expr, lam, want):
    assert _eval(expr, lam) == pytest.approx(want, rel=1e-15, abs=1e-15)


@pytest.mark.parametrize("bad", ["1.5+", "foo(2)", "x+1", "(1.5", "1.5)", ""])
def test_expression_errors(bad):
    with pytest.raises(ValueError):
        _eval(bad)


def test_expression_nesting_is_bounded():
    """Nesting is refused beyond a bound instead of overflowing the parser's
    stack (host-code sanitizer finding: 5000 nested parentheses); 400 levels
    of parentheses and of unary signs still evaluate."""
    assert _eval("(" * 400 + "1.5" + ")" * 400) == 1.5
    assert _eval("-" * 400 + "2") == 2.0
    for deep in ("(" * 5000 + "1" + ")" * 5000, "-" * 100000 + "1", "2^" * 3000 + "1"):
        with pytest.raises(ValueError, match="nested too deeply"):
            _eval(deep)


def _small():
    s = fields.make_settings((8, 6, 3), (0.1, 0.1, 0.1), [0.6], no="1.5", ne="1.7")
    lc = np.zeros((3, 6, 8, 3)); lc[..., 2] = 1
    E = np.zeros(fields.output_shape(s), dtype=np.complex128)
    return s, lc, E


def _run_strict(s, lc, E, mask=None, formula=""):
    js = s if isinstance(s, str) else json.dumps(s)
    if mask is None:
        bpm_backend.run_backend_without_mask(js, lc, E, strict=True)
    else:
        bpm_backend.run_backend_with_mask(js, formula, lc, mask, E, strict=True)


@pytest.mark.parametrize("mutate,msg", [
    (lambda s: s.pop("Algorithm settings"), 'The section "Algorithm settings" is missing from the setting file'),
    (lambda s: s["Physics settings"]["Coefficients"].pop("ne"), 'Error when parsing variable "ne" in section "Coefficients"'),
    (lambda s: s["Algorithm settings"]["General"].__setitem__("LC field type", "Nematic"), '"LC field type" must be'),
    (lambda s: s["Physics settings"]["Initial conditions"].__setitem__("Basis convention", "ABC"), '"Basis convention" must be'),
    (lambda s: s["Physics settings"]["Initial conditions"].__setitem__("Beam profile", "Gauss(2)"), "Unrecognised beam profile"),
    (lambda s: s["Physics settings"]["Coefficients"].__setitem__("Wavevectors", [0.0, 0.0, 0.1]), 'multiple of 2'),
    (lambda s: s["Physics settings"]["Coefficients"].__setitem__("no", "1.5*("), "Wrong expression for no"),
    (lambda s: s["Postprocessor settings"]["Screen output"].__setitem__("Isotropic layer thicknesses", [1.0]),
     "Inconsistent number of isotropic layers"),
    (lambda s: s["Physics settings"]["Coefficients"].__setitem__("no", 1.5), 'type must be string, but is number'),
])
def test_settings_errors_mirror_reference(mutate, msg):
    s, lc, E = _small()
    mutate(s)
    with pytest.raises(bpm_backend.BpmBackendError, match=re.escape(msg)):
        _run_strict(s, lc, E)


def test_array_size_errors():
    s, lc, E = _small()
    with pytest.raises(bpm_backend.BpmBackendError, match=r"Wrong dimension for the user pointer array \(VectorField\)"):
        _run_strict(s, lc.reshape(-1)[:-1], E)
    with pytest.raises(bpm_backend.BpmBackendError, match="Unexpected dimension for the LC orientational field"):
        _run_strict(s, np.zeros(8 * 6 * 3 * 4), E)
    E2 = np.zeros(E.size - 1, dtype=np.complex128)
    with pytest.raises(bpm_backend.BpmBackendError, match="OpticalFieldCollection"):
        _run_strict(s, lc, E2)
    with pytest.raises(bpm_backend.BpmBackendError, match="Wrong dimension for the mask pointer array"):
        _run_strict(s, lc, E, mask=np.ones(5))


def test_swig_array_typemaps():
    s, lc, E = _small()
    js = json.dumps(s)
    with pytest.raises(TypeError):
        bpm_backend.run_backend_without_mask(js, lc, E.astype(np.complex64))
    with pytest.raises(TypeError):
        bpm_backend.run_backend_without_mask(js, lc, np.zeros((2,) + E.shape, np.complex128)[:, ...][::2][0][..., ::1].T)


def test_default_error_behaviour_prints_and_returns(capsys):
    s, lc, E = _small()
    s["Physics settings"]["Initial conditions"]["Beam profile"] = "Laser"
    rv = bpm_backend.run_backend_without_mask(json.dumps(s), lc, E, strict=False)
    assert rv is None
    err = capsys.readouterr().err
    assert "Exception on processing" in err and "Unrecognised beam profile" in err and "Aborting!" in err


def test_wavelength_and_wavevector_generators_follow_reference():
    # SET/PhysicsSettings.cpp:43-82: spectral spread and condenser rings
    s, lc, E = _small()
    co = s["Physics settings"]["Coefficients"]
    del co["Wavelengths"], co["Wavevectors"]
    co.update({"Mean wavelength": 0.6, "Spectral FWHM": 0.2, "N wavelengths": 3,
               "Condenser numerical aperture": 0.2, "N radial illumination directions": 2})
    # sizes: 3 wavelengths x 7 wavevectors; a wrong-size output array is rejected
    E_bad = np.zeros((3, 6, 4, 6, 8), np.complex128)
    with pytest.raises(bpm_backend.BpmBackendError, match="OpticalFieldCollection"):
        _run_strict(s, lc, E_bad)


def test_gpu_path_fails_loudly_without_device():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    s, lc, E = _small()
    with pytest.raises(bpm_backend.BpmBackendError, match="HIP"):
        _run_strict(s, lc, E)


def test_footprint_estimate_scales():
    lib = _lib.load()
    s = fields.make_settings((512, 512, 100), (0.1,) * 3, [0.55], no="1.52", ne="1.74")
    out = ctypes.c_int64()
    err = _lib.err_buf()
    o64, o32 = _lib.options(precision=64), _lib.options(precision=32)
    assert lib.nmk_bpm_estimate_footprint(json.dumps(s).encode(), 512 * 512 * 100 * 3, 0, ctypes.byref(o64),
                                          ctypes.byref(out), err, len(err)) == 0
    b64 = out.value
    assert lib.nmk_bpm_estimate_footprint(json.dumps(s).encode(), 512 * 512 * 100 * 3, 0, ctypes.byref(o32),
                                          ctypes.byref(out), err, len(err)) == 0
    assert 512 * 512 * 100 * 24 < out.value < b64


@pytest.mark.parametrize("expr,lam", [("1.7+0.01/lambda^2", 0.55), ("-2^2", 0.6), ("2^3^2", 0.6),
                                      ("1.5-0.003/lambda^2+sqrt(lambda)", 0.45)])
def test_oracle_expression_evaluator_agrees_with_backend(expr, lam):
    import oracle
    assert oracle.eval_expr(expr, **{"lambda": lam}) == pytest.approx(_eval(expr, lam), rel=1e-15)


# --------------------------------------------------------------------------- mask-formula weights (host side)
def _host_weights(s, formula, mask):
    Nx, Ny, Nz = s["Physics settings"]["Initial conditions"]["Mesh dimensions"]
    w = np.empty(2 * Nx * Ny * (Nz - 1))
    err = _lib.err_buf()
    m = np.ascontiguousarray(mask, dtype=np.float64).reshape(-1)
    rc = _lib.load().nmk_bpm_interface_weights(json.dumps(s).encode(), formula.encode(), _lib.dptr(m), m.size,
                                               _lib.dptr(w), w.size, err, len(err))
    if rc:
        raise ValueError(err.value.decode())
    return w.reshape(Nz - 1, Ny, Nx, 2), err.value.decode()


def test_backend_interface_weights_closed_form_and_oracle():
    """The weights the backend computes from a mask formula (host code,
    UtilitiesLib/VectorField.cpp:64-113) against the closed-form crossings of a
    droplet R^2 - r^2 (within the 1e-3 dz bisection bracket) and against the
    oracle's independent numpy bisection (same steps, so equal to rounding)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mask_weights
    dims, sp, R = (13, 12, 10), (0.3, 0.3, 0.3), 1.5
    s = fields.make_settings(dims, sp, [0.6])
    x, y, z = fields.mesh_coords(*dims, *sp)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    mask = R * R - X * X - Y * Y - Z * Z
    w, notice = _host_weights(s, "%r^2-x^2-y^2-z^2" % R, mask)
    assert notice == ""
    w_or, bad = mask_weights.interface_weights(lambda a, b, c: R ** 2 - a ** 2 - b ** 2 - c ** 2, dims, sp, mask)
    assert not bad
    np.testing.assert_allclose(w, w_or, rtol=0, atol=1e-12)
    rho2 = R * R - x[None, :] ** 2 - y[:, None] ** 2
    zi = (np.arange(dims[2] - 1) - 0.5 * (dims[2] - 1)) * sp[2]
    for iz in range(dims[2] - 1):
        for sgn in (1.0, -1.0):
            with np.errstate(invalid="ignore"):
                root = sgn * np.sqrt(rho2)
            inside = (rho2 > 0) & (root > zi[iz]) & (root < zi[iz] + sp[2])
            assert np.all(np.abs(w[iz, ..., 0][inside] - (root[inside] - zi[iz]) / sp[2]) < 1e-3)
    assert np.sum(w[..., 0] != 0.5) > 20


def test_backend_interface_weights_mismatch_notice():
    """A formula that disagrees with the mask: all weights 0.5 and the
    reference's notice (UtilitiesLib/VectorField.cpp:107-111) with status 0."""
    dims, sp = (6, 5, 6), (0.2, 0.2, 0.2)
    s = fields.make_settings(dims, sp, [0.6])
    w, notice = _host_weights(s, "0.05-z", np.ones(dims[::-1]))
    assert np.all(w == 0.5)
    assert notice == "Mask values do not corresponds to the given formula, will ignore interpolation correction"
    with pytest.raises(ValueError, match="Wrong mask formula"):
        _host_weights(s, "0.05-", np.ones(dims[::-1]))


GRAMMAR_FORMULA = "(z>0.05 ? 0.6 : 0.7)^2-x^2-y^2-z^2+0*min(x,y)+0*avg(x,y,z)+0*abs(sqrt(x^2+1))"


def _grammar_fn(x, y, z):
    return (np.where(z > 0.05, 0.6, 0.7) ** 2 - x ** 2 - y ** 2 - z ** 2 + 0 * np.minimum(x, y) + 0 * ((x + y + z) / 3)
            + 0 * np.abs(np.sqrt(x ** 2 + 1)))


def test_host_weights_compiled_and_tree_evaluations_agree():
    """The mask formula is evaluated by its compiled postfix program (the one
    the device runs); a formula too long for a program falls back to the
    expression tree.  Both give the oracle's weights: a formula with a
    ternary, comparisons and functions (exact operations only, so equal to
    the bit), and a plane written short (compiled) and padded with 40 zero
    terms (over the program's size: tree)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mask_weights
    dims, sp = (14, 11, 9), (0.1, 0.1, 0.1)
    s = fields.make_settings(dims, sp, [0.6])
    x, y, z = fields.mesh_coords(*dims, *sp)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    mask = _grammar_fn(X, Y, Z)
    w, notice = _host_weights(s, GRAMMAR_FORMULA, mask)
    w_or, bad = mask_weights.interface_weights(_grammar_fn, dims, sp, mask)
    assert notice == "" and not bad
    assert np.array_equal(w, w_or)
    assert np.sum(w[..., 0] != 0.5) > 20
    plane = 0.13 - Z + 0 * X
    short, _ = _host_weights(s, "0.13-z", plane)
    long_, _ = _host_weights(s, "0.13-z" + "+0*x" * 40, plane)
    assert np.array_equal(short, long_) and np.sum(short[..., 0] != 0.5) == dims[0] * dims[1]


# --------------------------------------------------------------------------- the two drop-in C symbols
def test_drop_in_symbols_validate_before_any_device_work():
    """nmk_bpm_run_backend_without_mask / _with_mask called directly through
    ctypes (the binding INTEGRATION.md gives): input errors come back as
    NMK_BPM_EINPUT with the reference's message, before any device call."""
    lib = _lib.load()
    s, lc, E = _small()
    js = json.dumps(s).encode()
    lcf = np.ascontiguousarray(lc.reshape(-1))
    Ef = E.reshape(-1).view(np.float64)
    err = _lib.err_buf()
    rc = lib.nmk_bpm_run_backend_without_mask(js, _lib.dptr(lcf), lcf.size - 1, _lib.dptr(Ef), E.size, err, len(err))
    assert rc == 1 and b"Wrong dimension for the user pointer array" in err.value
    rc = lib.nmk_bpm_run_backend_without_mask(js, _lib.dptr(lcf), lcf.size, _lib.dptr(Ef), E.size - 1, err, len(err))
    assert rc == 1 and b"OpticalFieldCollection" in err.value
    m = np.ones(8 * 6 * 3)
    rc = lib.nmk_bpm_run_backend_with_mask(js, b"", _lib.dptr(lcf), lcf.size, _lib.dptr(m), m.size - 2,
                                           _lib.dptr(Ef), E.size, err, len(err))
    assert rc == 1 and b"Wrong dimension for the mask pointer array" in err.value
    rc = lib.nmk_bpm_run_backend_with_mask(js, b"", _lib.dptr(lcf), lcf.size, None, 0, _lib.dptr(Ef), E.size,
                                           err, len(err))
    assert rc == 1 and b"mask" in err.value


def test_one_shot_calls_refuse_graph_replay():
    s, lc, E = _small()
    with pytest.raises(ValueError, match="Context"):
        bpm_backend.run_backend_without_mask(json.dumps(s), lc, E, use_graph=True)


def test_noalloc_entry_point_refusals_before_device_work():
    """nmk_bpm_run_noalloc checks its inputs and the scratch size before any
    device call (the device pointers below are never dereferenced)."""
    lib = _lib.load()
    err = _lib.err_buf()
    fake = ctypes.c_void_p(1 << 20)
    s = fields.make_settings((1024, 1024, 3), (0.1,) * 3, [0.6])
    n_lc, n_E = 1024 * 1024 * 3 * 3, 4 * 1024 * 1024
    rc = lib.nmk_bpm_run_noalloc(json.dumps(s).encode(), fake, n_lc, None, 0, None, 0, fake, n_E, 64, None, err,
                                 len(err))
    assert rc == 3 and b"bytes of device scratch" in err.value, err.value
    rc = lib.nmk_bpm_run_noalloc(json.dumps(s).encode(), fake, n_lc - 1, None, 0, None, 0, fake, n_E, 64, None, err,
                                 len(err))
    assert rc == 1 and b"Wrong dimension for the user pointer array" in err.value
    rc = lib.nmk_bpm_run_noalloc(json.dumps(s).encode(), None, n_lc, None, 0, None, 0, fake, n_E, 64, None, err,
                                 len(err))
    assert rc == 1 and b"null device array" in err.value
