"""Microscope imaging of the bpm output on the device: the ``OpticalFields``
and ``FieldViewer`` classes of Nemaktis with their compute moved to HIP.

Mirrors ``HighLevelPythonInterface/nemaktis/light_propagator.py:414-846``
(``OpticalFields``: objective filter, ``update_NA_objective``,
``focus_fields``, ``focused_vals``) and
``HighLevelPythonInterface/nemaktis/field_viewer.py:16-517``
(``FieldViewer``: optical elements, ``update_image``, ``get_image``,
``get_spectrum``).  The refocus (2-D FFT, objective mask and defocus phase,
inverse FFT) and the micrograph synthesis (Jones matrices, q-integration,
colour conversion) run through the C ABI of ``include/nemaktis_view.h``; the
fields stay in HBM between calls and only the image (or what the caller asks
for) comes back to the host.  There is no CPU fallback.

Differences from the reference, by construction:
  * The GUI (``FieldViewer.plot``) is not provided; ``OpticalFields(vti_file=...)``
    and ``save_to_vti`` read and write the reference's file layout without VTK
    (``vti.py``).
  * Without ``cmf`` the colour table is ``colour.default_cmf`` (CIE 1931
    functions by a published analytic fit, a 6504 K black body for D65,
    normalised to the D65 white point) instead of dtmm's tabulated
    ``dtmm.color.load_tcmf``, which is not shipped here.
  * ``FieldViewer.NA_objective`` returns the objective NA (the reference
    declares its setter with ``@z_focus.setter``, so its getter returns the
    focus altitude).
"""
import ctypes
import os

import numpy as np

from . import _lib, colour, vti


class ViewerError(RuntimeError):
    pass


def _eval_index(expr, lam):
    """An index expression of the settings at wavelength lam (um), evaluated as the backend evaluates it."""
    lib = _lib.load()
    v = ctypes.c_double()
    err = _lib.err_buf()
    _check(lib.nmk_bpm_eval_expression(expr.encode(), float(lam), ctypes.byref(v), err, len(err)), err)
    return v.value


def _check(rc, err):
    if rc:
        raise ViewerError(err.value.decode())


def _geometry(Nx, Ny, Lx, Ly, wavelengths, n_out, wavevectors, max_NA_objective, max_NA_condenser, Nr,
              koehler_1D, zfoc_NA_corr, device):
    wl = np.ascontiguousarray(np.atleast_1d(np.asarray(wavelengths, dtype=np.float64)))
    nout = np.ascontiguousarray(np.broadcast_to(np.asarray(n_out, dtype=np.float64), wl.shape))
    g = _lib.ViewGeometry()
    g.Nx, g.Ny, g.Lx, g.Ly = int(Nx), int(Ny), float(Lx), float(Ly)
    g.n_wavelengths = wl.size
    g.wavelengths = _lib.dptr(wl)
    g.n_out = _lib.dptr(nout)
    keep = [wl, nout]
    if wavevectors is not None:
        q = np.ascontiguousarray(np.asarray(wavevectors, dtype=np.float64).reshape(-1, 2))
        g.n_wavevectors = q.shape[0]
        g.wavevectors = _lib.dptr(q)
        keep.append(q)
    g.max_NA_objective = float(max_NA_objective)
    g.max_NA_condenser = float(max_NA_condenser)
    g.N_radial_wavevectors = int(Nr)
    g.koehler_1D = 1 if koehler_1D else 0
    g.zfoc_NA_corr = float(zfoc_NA_corr)
    g.device = int(device)
    return g, keep


def condenser_wavevectors(mesh_dimensions, max_NA_condenser, N_radial_wavevectors, koehler_1D=False):
    """Wavevectors (qx, qy) of the condenser rings (light_propagator.py:578-599)."""
    Nr = N_radial_wavevectors
    if mesh_dimensions[0] == 1 and koehler_1D:
        q = np.zeros((2 * Nr - 1, 2))
        for ir in range(1, Nr):
            q[Nr - 1 + ir, 1] = ir * max_NA_condenser / (Nr - 1)
            q[Nr - 1 - ir, 1] = -ir * max_NA_condenser / (Nr - 1)
    elif mesh_dimensions[1] == 1 and koehler_1D:
        q = np.zeros((2 * Nr - 1, 2))
        for ir in range(1, Nr):
            q[Nr - 1 + ir, 0] = ir * max_NA_condenser / (Nr - 1)
            q[Nr - 1 - ir, 0] = -ir * max_NA_condenser / (Nr - 1)
    else:
        q = np.zeros((1 + 3 * Nr * (Nr - 1), 2))
        for ir in range(1, Nr):
            beta = ir * max_NA_condenser / (Nr - 1)
            for iphi in range(6 * ir):
                phi = iphi * np.pi / (3 * ir)
                q[1 + 3 * ir * (ir - 1) + iphi] = beta * np.cos(phi), beta * np.sin(phi)
    return q


class OpticalFields:
    """Transverse output fields of shape (N_wavelengths, N_wavevectors, 4, Ny, Nx) and their refocus
    through the microscope objective (reference: light_propagator.py:414-846).

    ``OpticalFields(wavelengths=..., max_NA_objective=..., mesh_lengths=(Lx, Ly), mesh_dimensions=(Nx, Ny),
    max_NA_condenser=0, N_radial_wavevectors=1, koehler_1D=False, zfoc_NA_corr=0, n_out=1, device=-1)``
    """

    def __init__(self, **kwargs):
        if len(kwargs) == 1 and "vti_file" in kwargs:
            self._init_from_vti(kwargs["vti_file"])
            return
        if len(kwargs["mesh_dimensions"]) != 2:
            raise Exception("mesh_dimensions should be an array-like object of size 2")
        if len(kwargs["mesh_lengths"]) != 2:
            raise Exception("mesh_lengths should be an array-like object of size 2")
        self._n_out = kwargs.get("n_out", 1)
        self._zfoc_NA_corr = kwargs.get("zfoc_NA_corr", 0)
        self._wavelengths = np.array(kwargs["wavelengths"], dtype=np.float64).reshape(-1)
        self._max_NA_objective = kwargs["max_NA_objective"]
        self._max_NA_condenser = kwargs.get("max_NA_condenser", 0)
        self._N_radial_wavevectors = kwargs.get("N_radial_wavevectors", 1)
        self._koehler_1D = kwargs.get("koehler_1D", False)
        self._NA = self._max_NA_objective
        self._device = kwargs.get("device", -1)
        dims = np.array(kwargs["mesh_dimensions"])
        lengths = np.array(kwargs["mesh_lengths"], dtype=np.float64)
        self._wavevectors = condenser_wavevectors(dims, self._max_NA_condenser, self._N_radial_wavevectors,
                                                  self._koehler_1D)
        (self._Nx, self._Ny) = (int(dims[0]), int(dims[1]))
        (self._Lx, self._Ly) = (float(lengths[0]), float(lengths[1]))
        (self._dx, self._dy) = tuple(lengths / np.maximum(np.ones(2), dims - 1))
        shape = (len(self._wavelengths), len(self._wavevectors), 4, self._Ny, self._Nx)
        self._vals = np.zeros(shape, dtype=np.complex128)
        self._init_common()

    def _init_from_vti(self, path):
        """Fields, wavelengths, wavevectors and objective data of a file written by ``save_to_vti`` (or by the
        reference's, light_propagator.py:449-541: the same array names, checks and messages), read without VTK."""
        if not os.path.isfile(path):
            raise Exception("VTI file does not exists")
        print("{ Initializing optical fields from " + path + " }")
        geo, fd, pd = vti.read_vti(path)
        dims, spacings = geo["dims"], geo["spacing"]
        if dims[2] != 1:
            raise Exception("The specified vti file should include 2D data")
        self._zfoc_NA_corr = float(np.ravel(fd["zfoc_NA_corr"])[0]) if "zfoc_NA_corr" in fd else 0
        if "lambda" not in fd:
            raise Exception("VTI file is missing the field array \"lambda\" for the wavelengths")
        self._wavelengths = np.asarray(fd["lambda"], dtype=np.float64).reshape(-1)
        Nl = len(self._wavelengths)
        if "n_out" in fd:
            n_out = np.asarray(fd["n_out"], dtype=np.float64).reshape(-1)
            self._n_out = float(n_out[0]) if n_out.size == 1 else n_out
        else:
            self._n_out = np.ones(Nl)
        for k in ("qx", "qy"):
            if k not in fd:
                raise Exception("VTI file is missing the field array \"%s\" for the wavevectors" % k)
        self._wavevectors = np.stack((np.asarray(fd["qx"], dtype=np.float64).reshape(-1),
                                      np.asarray(fd["qy"], dtype=np.float64).reshape(-1)), axis=-1)
        Nq = len(self._wavevectors)
        self._koehler_1D = bool(np.ravel(fd["koehler_1D"])[0] == 1) if "koehler_1D" in fd else False
        if (dims[0] == 1 or dims[1] == 1) and self._koehler_1D:
            Nr = Nq // 2
            if Nq != 1 + 2 * Nr:
                raise Exception("VTI file contain the wrong number of wavevectors")
        else:
            Nr = int(np.round((1 + np.sqrt((4 * Nq - 1) / 3)) / 2))
            if Nq != 1 + 3 * Nr * (Nr - 1):
                raise Exception("VTI file contain the wrong number of wavevectors")
        self._N_radial_wavevectors = Nr
        self._max_NA_condenser = float(np.sqrt(np.sum(self._wavevectors ** 2, axis=1))[-1])
        if "max_NA_objective" not in fd:
            raise Exception("VTI file is missing the field scalar \"max_NA_objective\"")
        self._max_NA_objective = float(np.ravel(fd["max_NA_objective"])[0])
        self._NA = self._max_NA_objective
        (self._Nx, self._Ny) = (int(dims[0]), int(dims[1]))
        (self._dx, self._dy) = (float(spacings[0]), float(spacings[1]))
        (self._Lx, self._Ly) = (self._dx * (self._Nx - 1), self._dy * (self._Ny - 1))
        self._device = -1
        vals = np.zeros((Nl, Nq, 4, self._Ny, self._Nx), dtype=np.complex128)
        for w in range(Nl):
            for q in range(Nq):
                names = ["E_%s_input%s_%d_%d" % (part, pol, w, q) for pol in "XY" for part in ("real", "imag")]
                for name in names:
                    if name not in pd:
                        raise Exception("Missing array \"" + name + "\" in VTI file")
                for k, pol in enumerate("XY"):
                    re_ = pd["E_real_input%s_%d_%d" % (pol, w, q)].reshape(self._Ny, self._Nx, -1)
                    im_ = pd["E_imag_input%s_%d_%d" % (pol, w, q)].reshape(self._Ny, self._Nx, -1)
                    vals[w, q, 2 * k] = re_[..., 0] + 1j * im_[..., 0]
                    vals[w, q, 2 * k + 1] = re_[..., 1] + 1j * im_[..., 1]
        self._vals = vals
        self._init_common()

    def save_to_vti(self, filename):
        """Writes the fields and the data needed to read them back (wavelengths, wavevectors, n_out, focus
        correction, Koehler flag, objective NA) as a VTK image-data file (reference:
        light_propagator.py:722-790; uncompressed raw appended arrays, readable by VTK and ``vti.read_vti``);
        ".vti" is appended unless the name ends with it."""
        path = filename if filename[-4:] == ".vti" else filename + ".vti"
        print("{ Saving optical fields to " + path + " }")
        vals = self.vals
        Np = self._Nx * self._Ny
        point = {}
        for w in range(len(self._wavelengths)):
            for q in range(len(self._wavevectors)):
                for k, pol in enumerate("XY"):
                    E = vals[w, q, [2 * k, 2 * k + 1], :, :].reshape(2, Np).T
                    point["E_real_input%s_%d_%d" % (pol, w, q)] = np.real(E)
                    point["E_imag_input%s_%d_%d" % (pol, w, q)] = np.imag(E)
        field = {"lambda": self._wavelengths, "zfoc_NA_corr": [self.zfoc_NA_corr],
                 "n_out": np.atleast_1d(np.asarray(self.n_out, dtype=np.float64)),
                 "koehler_1D": [1.0 if self._koehler_1D else 0.0], "qx": self._wavevectors[:, 0],
                 "qy": self._wavevectors[:, 1], "max_NA_objective": [self._max_NA_objective]}
        vti.write_vti(path, (self._Nx, self._Ny, 1), (self._dx, self._dy, 0.0), (-self._Lx / 2, -self._Ly / 2, 0.0),
                      field, point)

    def _init_common(self):
        if self._N_radial_wavevectors > 1 and self._max_NA_condenser > 0:
            self._delta_qr = self._max_NA_condenser / (self._N_radial_wavevectors - 1)
        else:
            self._delta_qr = 1
        self._z = 0
        self._lib = _lib.load()
        self._h = None
        self._ctx = None
        self._upload = True
        self._focused = None

    @classmethod
    def from_context(cls, ctx):
        """Fields of the last ``propagate()`` of a ``bpm_backend.Context``, copied on the device: the settings
        give the mesh, wavelengths, n_out, wavevectors, objective NA and focus correction as
        ``LightPropagator._bpm_propagation`` passes them (light_propagator.py:196-219)."""
        self = cls.__new__(cls)
        lib = _lib.load()
        err = _lib.err_buf()
        h = lib.nmk_view_from_ctx(ctx.h, None, err, len(err))
        if not h:
            raise ViewerError(err.value.decode())
        import json
        s = json.loads(ctx.json_str)
        ic = s["Physics settings"]["Initial conditions"]
        co = s["Physics settings"]["Coefficients"]
        scr = s["Postprocessor settings"]["Screen output"]
        dims, sp = ic["Mesh dimensions"], ic["Mesh spacings"]
        self._wavelengths = np.array(co["Wavelengths"], dtype=np.float64).reshape(-1)
        self._wavevectors = np.array(co["Wavevectors"], dtype=np.float64).reshape(-1, 2)
        self._Nx, self._Ny = int(dims[0]), int(dims[1])
        self._dx, self._dy = float(sp[0]), float(sp[1])
        self._Lx, self._Ly = self._dx * max(1, self._Nx - 1), self._dy * max(1, self._Ny - 1)
        self._max_NA_objective = float(scr["Numerical aperture"])
        self._NA = self._max_NA_objective
        Nq = len(self._wavevectors)
        self._N_radial_wavevectors = int(np.round((1 + np.sqrt((4 * Nq - 1) / 3)) / 2))
        self._max_NA_condenser = float(np.sqrt(np.sum(self._wavevectors ** 2, axis=1))[-1])
        self._koehler_1D = False
        # n_out per wavelength and the focus correction of the isotropic layers, as the device view has them
        # (light_propagator.py:194-207)
        nout = str(co.get("nout", "1"))
        self._n_out = np.array([_eval_index(nout, lam) for lam in self._wavelengths])
        n0 = _eval_index(nout, 0.6)
        self._zfoc_NA_corr = 0.0
        for ek, nk in zip(scr.get("Isotropic layer thicknesses", []), scr.get("Isotropic layer refractive indices", [])):
            self._zfoc_NA_corr += 3 * n0 * ek / (14 * nk) * (1 / n0 ** 2 - 1 / nk ** 2)
        self._device = -1
        self._vals = None
        self._init_common()
        self._h = h
        self._ctx = ctx
        self._upload = False
        return self

    # ---- device handle -------------------------------------------------------------------------------------
    def _handle(self):
        if self._h is None:
            g, keep = _geometry(self._Nx, self._Ny, self._Lx, self._Ly, self._wavelengths, self._n_out,
                                self._wavevectors, self._max_NA_objective, self._max_NA_condenser,
                                self._N_radial_wavevectors, self._koehler_1D, self._zfoc_NA_corr, self._device)
            err = _lib.err_buf()
            self._h = self._lib.nmk_view_create(ctypes.byref(g), None, 0, err, len(err))
            del keep
            if not self._h:
                raise ViewerError(err.value.decode())
            self._upload = True
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nmk_view_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference interface -------------------------------------------------------------------------------
    def copy(self):
        """Independent object with the same geometry and its own copy of the field values."""
        new_fields = OpticalFields(
            wavelengths=self._wavelengths, max_NA_objective=self._max_NA_objective,
            max_NA_condenser=self._max_NA_condenser, N_radial_wavevectors=self._N_radial_wavevectors,
            koehler_1D=self._koehler_1D, zfoc_NA_corr=self.zfoc_NA_corr, n_out=self.n_out,
            mesh_dimensions=(self._Nx, self._Ny), mesh_lengths=(self._Lx, self._Ly), device=self._device)
        new_fields.vals = self.vals
        return new_fields

    @property
    def n_out(self):
        if self._n_out is None:
            raise ViewerError("n_out is held on the device for fields taken from a context")
        return self._n_out

    @property
    def zfoc_NA_corr(self):
        if self._zfoc_NA_corr is None:
            raise ViewerError("zfoc_NA_corr is held on the device for fields taken from a context")
        return self._zfoc_NA_corr

    @property
    def focused_vals(self):
        """Output of the last focus_fields(): complex128 array of the same shape as ``vals``."""
        if self._focused is None:
            out = np.empty(self._shape(), dtype=np.complex128)
            err = _lib.err_buf()
            _check(self._lib.nmk_view_get_focused(self._handle(), _lib.dptr(out.reshape(-1).view(np.float64)),
                                                  out.size, err, len(err)), err)
            self._focused = out
        return self._focused

    def _shape(self):
        return (len(self._wavelengths), len(self._wavevectors), 4, self._Ny, self._Nx)

    @property
    def vals(self):
        """Transverse fields, complex128 (N_wavelengths, N_wavevectors, 4, Ny, Nx): the last axis-2 index
        runs over (Ex, Ey) of the x-polarised input, then (Ex, Ey) of the y-polarised input."""
        if self._vals is None:
            self._vals = np.empty(self._shape(), dtype=np.complex128)
            self._ctx.download(self._vals)
        # the caller may write into the array: upload it again before the next focus
        self._upload = True
        return self._vals

    @vals.setter
    def vals(self, fields_ndarray):
        if self._shape() == fields_ndarray.shape:
            if self._vals is None:
                self._vals = np.empty(self._shape(), dtype=np.complex128)
            self._vals[:] = fields_ndarray
            self._upload = True
        else:
            raise Exception("Wrong shape for the optical field ndarray")

    @property
    def z_focus(self):
        return self._z

    def update_NA_objective(self, new_NA):
        self._NA = max(0, min(self._max_NA_objective, new_NA))

    def focus_fields(self, z_focus=None):
        """Refocuses the fields: the objective filter for focal height z_focus (kept for later calls when
        omitted) and the current NA, applied to every (wavelength, wavevector) plane on the device."""
        if z_focus is not None:
            self._z = z_focus
        h = self._handle()
        err = _lib.err_buf()
        if self._upload:
            v = np.ascontiguousarray(self._vals, dtype=np.complex128)
            _check(self._lib.nmk_view_set_fields(h, _lib.dptr(v.reshape(-1).view(np.float64)), v.size, err,
                                                 len(err)), err)
            self._upload = False
        _check(self._lib.nmk_view_focus(h, float(self._z), float(self._NA), err, len(err)), err)
        self._focused = None

    def timings(self):
        """Device ms of the last focus / render: filter, rows forward, columns, rows backward, spectrum, colour."""
        ms = (ctypes.c_double * 6)()
        self._lib.nmk_view_timings(self._handle(), ms, 6)
        return dict(zip(["filter", "rows_fwd", "columns", "rows_bwd", "spectrum", "colour"], list(ms)))

    def get_pos(self, ix, iy):
        return np.array([(ix - (self._Nx - 1) / 2) * self._dx, (iy - (self._Ny - 1) / 2) * self._dy])

    def get_wavelengths(self):
        return self._wavelengths

    def get_wavevectors(self):
        return self._wavevectors

    def get_qr_index(self, NA_condenser):
        return int(np.ceil(NA_condenser / self._delta_qr))

    def get_delta_qr(self):
        return self._delta_qr

    def get_mesh_dimensions(self):
        return (self._Nx, self._Ny)

    def get_mesh_lengths(self):
        return (self._Lx, self._Ly)

    def get_mesh_spacings(self):
        return (self._dx, self._dy)

    def get_n_vertices(self):
        return self._Nx * self._Ny


class FieldViewer:
    """Turns the transverse fields into polarised-light micrographs on the device: polariser, waveplates and
    analyser applied per pixel, incoherent sum over the condenser wavevectors, colour from the spectrum
    (reference: field_viewer.py:16-470).

    ``FieldViewer(optical_fields, cmf=None, gamma=True)``; ``cmf`` is the (N_wavelengths, 3) colour matching
    table of the wavelengths (the ``cmf`` argument of the reference); ``None`` selects
    ``colour.default_cmf`` of the fields' wavelengths."""

    polariser = True
    analyser = True
    upper_waveplate = "No"
    lower_waveplate = "No"
    polariser_angle = 0
    analyser_angle = 90
    upper_waveplate_angle = 0
    lower_waveplate_angle = 0
    angle_lock = False
    intensity = 1
    NA_condenser = 0
    n_tiles_x = 1
    n_tiles_y = 1
    grayscale = False

    def __init__(self, optical_fields, cmf=None, gamma=True):
        if not isinstance(optical_fields, OpticalFields):
            raise TypeError("optical_fields should be an OpticalFields object")
        self._optical_fields = optical_fields
        optical_fields.focus_fields(0)
        self._NA_objective = optical_fields._max_NA_objective
        self._gamma = gamma
        Nl = len(optical_fields.get_wavelengths())
        Nx, Ny = optical_fields.get_mesh_dimensions()
        self._q_averaged_specter = np.zeros((Ny, Nx, Nl))
        self._image = np.zeros((Ny, Nx, 3))
        if cmf is None:
            cmf = colour.default_cmf(optical_fields.get_wavelengths())
        self._cmf = np.ascontiguousarray(np.asarray(cmf, dtype=np.float64).reshape(Nl, 3))
        self._koehler_1D = optical_fields._koehler_1D
        self._N_radial_wavevectors = optical_fields._N_radial_wavevectors

    def plot(self):
        raise NotImplementedError("the graphical interface is not provided by nemaktis_amd")

    @property
    def z_focus(self):
        """Height z of the plane the objective is focused on."""
        return self._optical_fields.z_focus

    @z_focus.setter
    def z_focus(self, new_z):
        self._optical_fields.focus_fields(new_z)

    @property
    def NA_objective(self):
        """Objective NA used for the next render (clipped to the fields' maximum objective NA)."""
        return self._NA_objective

    @NA_objective.setter
    def NA_objective(self, new_NA):
        self._NA_objective = new_NA
        self._optical_fields.update_NA_objective(new_NA)
        self._optical_fields.focus_fields()

    def _optics(self):
        o = _lib.ViewOptics()
        lib = _lib.load()
        lib.nmk_view_default_optics(ctypes.byref(o))
        o.polariser = 1 if self.polariser else 0
        o.analyser = 1 if self.analyser else 0
        o.polariser_angle = float(self.polariser_angle)
        o.analyser_angle = float(self.analyser_angle)
        if self.lower_waveplate not in _lib.WAVEPLATES or self.upper_waveplate not in _lib.WAVEPLATES:
            raise ValueError("waveplates are 'No', 'Quarter-wave', 'Half-wave' or 'Tint-sensitive'")
        o.lower_waveplate = _lib.WAVEPLATES[self.lower_waveplate]
        o.upper_waveplate = _lib.WAVEPLATES[self.upper_waveplate]
        o.lower_waveplate_angle = float(self.lower_waveplate_angle)
        o.upper_waveplate_angle = float(self.upper_waveplate_angle)
        o.intensity = float(self.intensity)
        o.NA_condenser = float(self.NA_condenser)
        o.grayscale = 1 if self.grayscale else 0
        o.gamma = 1 if self._gamma else 0
        return o

    def render(self):
        """Computes spectrum (and image when a colour table is set) on the device without downloading."""
        of = self._optical_fields
        o = self._optics()
        err = _lib.err_buf()
        cmf = self._cmf
        _check(of._lib.nmk_view_render(of._handle(), ctypes.byref(o), _lib.dptr(cmf),
                                       0 if cmf is None else cmf.size, err, len(err)), err)

    def update_image(self):
        """Refocuses the fields if needed, then recomputes the spectrum and the image on the device."""
        self.render()
        of = self._optical_fields
        err = _lib.err_buf()
        _check(of._lib.nmk_view_get_spectrum(of._handle(), _lib.dptr(self._q_averaged_specter),
                                             self._q_averaged_specter.size, err, len(err)), err)
        if self._cmf is not None:
            _check(of._lib.nmk_view_get_image(of._handle(), _lib.dptr(self._image), self._image.size, err,
                                              len(err)), err)

    def get_image(self):
        """Last rendered image: float64 array (Ny, Nx, 3); needs a colour table."""
        if self._cmf is None:
            raise ViewerError("a colour matching table (cmf) is required for images")
        return np.flip(np.tile(self._image, (self.n_tiles_y, self.n_tiles_x, 1)), axis=0)

    def get_spectrum(self):
        """Last rendered spectrum, summed over the condenser wavevectors: float64 array (Ny, Nx, N_wavelengths)."""
        return np.flip(np.tile(self._q_averaged_specter, (self.n_tiles_y, self.n_tiles_x, 1)), axis=0)
