"""GPU parity of the imaging step (objective refocus and micrograph synthesis)
against the FP64 numpy oracle, through the C ABI of include/nemaktis_view.h.

Tolerances: refocused fields, relative L2 per plane <= 1e-12; spectra and
images, max abs <= 1e-12 relative to the largest value.  The device path
computes in FP64 like the reference (complex128 FFTs).
"""
import json

import numpy as np
import pytest

import viewer_oracle as vo
from nemaktis_amd import bpm_backend, colour, fields, viewer

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _rel(a, b):
    a = a.reshape(-1, a.shape[-2] * a.shape[-1])
    b = b.reshape(-1, b.shape[-2] * b.shape[-1])
    return float((np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max())


CASES = [
    # Nx, Ny, Lx, Ly, wavelengths, n_out, NA_obj, NA_cond, Nr, koehler, zfoc_corr
    (30, 24, 4.35, 3.45, [0.5, 0.6], [1.0, 1.33], 0.4, 0.2, 2, False, 0.0),
    (33, 49, 3.2, 4.8, [0.4], 1.0, 0.95, 0.0, 1, False, 0.0),        # radices 3, 11, 7; theta < -pi inside NA
    (52, 20, 5.1, 1.9, [0.45, 0.55, 0.65], 1.2, 0.6, 0.3, 3, False, -1.7),   # radix 13, focus correction
    (1, 40, 0.1, 3.9, [0.55, 0.6], 1.0, 0.5, 0.3, 3, True, 0.0),      # 1-D Koehler along y
    (256, 256, 25.5, 25.5, [0.6], 1.0, 0.4, 0.0, 1, False, 0.0),     # 8 columns per block
    (520, 64, 51.9, 6.3, [0.6], 1.0, 0.4, 0.0, 1, False, 0.0),       # rows of 520, columns of 64
    (17, 34, 1.6, 3.3, [0.55, 0.6], 1.0, 0.5, 0.2, 2, False, 0.0),    # prime 17: direct DFT rows and columns
    (97, 2500, 9.6, 249.9, [0.6], 1.0, 0.4, 0.0, 1, False, 0.0),     # prime 97, 2500 > 2048: direct DFT
]


def _make(case, seed=0):
    Nx, Ny, Lx, Ly, wl, nout, NA, NAc, Nr, koehler, zc = case
    of = viewer.OpticalFields(wavelengths=wl, max_NA_objective=NA, max_NA_condenser=NAc, N_radial_wavevectors=Nr,
                              koehler_1D=koehler, zfoc_NA_corr=zc, n_out=nout, mesh_lengths=(Lx, Ly),
                              mesh_dimensions=(Nx, Ny))
    of.vals = _fields(of.vals.shape, seed)
    return of


def _oracle_focus(of, z, NA):
    Nl = len(of.get_wavelengths())
    nout = np.broadcast_to(np.asarray(of._n_out, dtype=np.float64), (Nl,))
    dx, dy = of.get_mesh_spacings()
    return vo.focus(of.vals, dx, dy, of.get_wavelengths(), nout, of.get_wavevectors(), NA, z, of._zfoc_NA_corr)


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_focus_matches_oracle(ci):
    of = _make(CASES[ci], seed=ci)
    print("[gpu case] focus", CASES[ci][:2], flush=True)
    for z, NA in [(0.0, None), (1.7, None), (-2.35, 0.3), (4.0, 0.9)]:
        if NA is not None:
            of.update_NA_objective(NA)
        of.focus_fields(z)
        got = of.focused_vals
        want = _oracle_focus(of, z, of._NA)
        nz = np.linalg.norm(want.reshape(-1, want.shape[-2] * want.shape[-1]), axis=1) > 0
        assert nz.any()
        err = _rel(got.reshape(-1, *got.shape[-2:])[nz], want.reshape(-1, *want.shape[-2:])[nz])
        assert err < TOL, (z, NA, err)
        assert np.all(got.reshape(-1, *got.shape[-2:])[~nz] == 0)


OPTICS = [
    dict(),
    dict(polariser=False, analyser=False),
    dict(analyser_angle=30, polariser_angle=-10),
    dict(analyser=False, lower_waveplate="Quarter-wave", lower_waveplate_angle=20),
    dict(polariser=False, lower_waveplate="Half-wave", lower_waveplate_angle=35),
    dict(upper_waveplate="Tint-sensitive", upper_waveplate_angle=45),
    dict(lower_waveplate="Tint-sensitive", lower_waveplate_angle=10, upper_waveplate="Quarter-wave",
         upper_waveplate_angle=-45, analyser_angle=80),
]


@pytest.mark.parametrize("ci", [0, 2, 3])
def test_spectrum_and_image_match_oracle(ci):
    case = CASES[ci]
    of = _make(case, seed=10 + ci)
    Nl = len(of.get_wavelengths())
    rng = np.random.default_rng(ci)
    cmf = rng.uniform(0, 1, (Nl, 3))
    fv = viewer.FieldViewer(of, cmf=cmf)
    of.update_NA_objective(0.35)
    of.focus_fields(0.8)
    foc = of.focused_vals
    Nr, koehler, NAc_max = case[8], case[9], case[7]
    print("[gpu case] render", case[:2], flush=True)
    for k, optics in enumerate(OPTICS):
        for NAc in sorted({0.0, 0.5 * NAc_max, NAc_max, 0.37 * NAc_max}):
            for name, val in optics.items():
                setattr(fv, name, val)
            fv.NA_condenser = NAc
            fv.intensity = 1.3
            fv.grayscale = k == 3
            fv.update_image()
            opt = dict(optics, NA_condenser=NAc)
            want = vo.spectrum(foc, of.get_wavelengths(), opt, Nr, NAc_max, koehler)
            got = np.flip(fv.get_spectrum(), axis=0)
            scale = np.abs(want).max()
            np.testing.assert_allclose(got, want, rtol=0, atol=TOL * scale)
            img = vo.colour(want, cmf, 1.3, grayscale=fv.grayscale, gamma=True)
            np.testing.assert_allclose(np.flip(fv.get_image(), axis=0), img, rtol=0, atol=1e-11)
            for name in optics:
                setattr(fv, name, getattr(viewer.FieldViewer, name))


def test_image_without_gamma_is_clipped_and_tiles_flip():
    of = _make(CASES[0], seed=3)
    Nl = len(of.get_wavelengths())
    cmf = np.tile([[0.4, 0.5, 0.3]], (Nl, 1))
    fv = viewer.FieldViewer(of, cmf=cmf, gamma=False)
    fv.analyser = False
    fv.intensity = 3.0
    fv.n_tiles_x, fv.n_tiles_y = 2, 3
    fv.update_image()
    img = fv.get_image()
    Nx, Ny = of.get_mesh_dimensions()
    assert img.shape == (3 * Ny, 2 * Nx, 3)
    want = vo.colour(vo.spectrum(of.focused_vals, of.get_wavelengths(), dict(analyser=False), 2, 0.2), cmf, 3.0,
                     gamma=False)
    np.testing.assert_allclose(img, np.flip(np.tile(want, (3, 2, 1)), axis=0), atol=1e-12)
    assert img.min() >= 0 and img.max() <= 1


def test_view_from_context_equals_view_from_host_fields(tmp_path):
    s, lc, _, _ = fields.workload("tn64")
    n_E = len(s["Physics settings"]["Coefficients"]["Wavelengths"]) * 1 * 4 * 64 * 64
    with bpm_backend.Context(json.dumps(s), lc, n_E) as ctx:
        ctx.propagate()
        E = np.empty((1, 1, 4, 64, 64), dtype=np.complex128)
        ctx.download(E)
        dev = viewer.OpticalFields.from_context(ctx)
        host = viewer.OpticalFields(wavelengths=s["Physics settings"]["Coefficients"]["Wavelengths"],
                                    max_NA_objective=dev._max_NA_objective, mesh_dimensions=(64, 64),
                                    mesh_lengths=dev.get_mesh_lengths(), n_out=1.0)
        host.vals = E
        for of in (dev, host):
            of.focus_fields(1.2)
        np.testing.assert_array_equal(dev.focused_vals, host.focused_vals)
        cmf = np.array([[0.3, 0.6, 0.2]])
        fd, fh = viewer.FieldViewer(dev, cmf=cmf), viewer.FieldViewer(host, cmf=cmf)
        fd.update_image()
        fh.update_image()
        np.testing.assert_array_equal(fd.get_image(), fh.get_image())
        assert dev.vals.shape == E.shape and np.array_equal(dev.vals, E)
        # the fields taken from the context saved and read back as the reference's VTI file
        np.testing.assert_array_equal(dev.n_out, [1.0])
        dev.save_to_vti(str(tmp_path / "fields"))
        back = viewer.OpticalFields(vti_file=str(tmp_path / "fields.vti"))
        assert np.array_equal(back.vals, E) and back.zfoc_NA_corr == dev.zfoc_NA_corr == 0
        back.focus_fields(1.2)
        # the file's spacing against lengths / (N - 1): the filter may differ in the last place
        np.testing.assert_allclose(back.focused_vals, host.focused_vals, rtol=0, atol=1e-14)


def test_view_errors_are_reported():
    of = _make(CASES[0])
    fv = viewer.FieldViewer(of)
    fv.lower_waveplate = "Full-wave"
    with pytest.raises(ValueError):
        fv.update_image()


def test_image_with_the_default_colour_table():
    """Without cmf the viewer renders with colour.default_cmf of the fields'
    wavelengths (the table itself is pinned on the CPU); a fully
    transmitting medium (uniform x-polarised field, no polariser or
    analyser) renders as white."""
    of = _make(CASES[2], seed=5)
    wl = of.get_wavelengths()
    fv = viewer.FieldViewer(of)
    fv.analyser = False
    fv.update_image()
    want = vo.colour(vo.spectrum(of.focused_vals, wl, dict(analyser=False), CASES[2][8], CASES[2][7]),
                     colour.default_cmf(wl), 1.0, gamma=True)
    np.testing.assert_allclose(np.flip(fv.get_image(), axis=0), want, rtol=0, atol=1e-11)
    flat = _make(CASES[0], seed=6)
    v = np.zeros(flat.vals.shape, np.complex128)
    v[:, :, 0] = 1.0  # x input -> unit Ex
    v[:, :, 3] = 1.0  # y input -> unit Ey
    flat.vals = v
    fw = viewer.FieldViewer(flat)
    fw.polariser = fw.analyser = False
    fw.update_image()
    # D65 white through the sRGB matrix (7 significant digits) and the transfer function
    np.testing.assert_allclose(fw.get_image(), 1.0, atol=1e-5)
