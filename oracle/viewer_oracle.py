"""TEST INFRASTRUCTURE - NOT PRODUCT CODE.

FP64 numpy restatement of the microscope imaging step that Nemaktis runs
after the bpm backend.  Only tests/, the smoke check and bench.py's
cpu_baseline leg use it, as the checker or the CPU baseline.

  focus()           OpticalFields objective filter and focus_fields
                    (HighLevelPythonInterface/nemaktis/light_propagator.py:578-623, 690-719)
  quadrature()      q-integration weights of FieldViewer._update_averaged_specter
                    (HighLevelPythonInterface/nemaktis/field_viewer.py:388-446)
  jones()           polariser / waveplate matrices and their order in
                    FieldViewer._update_specter (field_viewer.py:347-385, 472-512)
  spectrum()        q-averaged spectrum (field_viewer.py:347-446)
  colour()          spectrum -> image of FieldViewer._update_POM_image
                    (field_viewer.py:449-455) through dtmm.color.specter2color

Pins (tests/test_viewer_host.py):
  focus       a single Fourier mode comes out multiplied by the closed-form
              filter value; refocusing at -z undoes +z on band-limited fields;
              zf = 0 with a wide aperture is the identity on band-limited
              fields; the principal-branch power equals numpy's complex power.
  quadrature  the weights integrate a q-independent spectrum exactly:
              2-D rings give s * (-log(1 - NA^2)) / NA^2, 1-D Koehler gives s.
  jones       Malus's law, crossed polarisers, half- and quarter-wave plates at
              45 degrees, tint plate at and off 540 nm (closed forms).
  colour      the sRGB transfer function at its ends and its junction, the D65
              white point mapped to (1, 1, 1), grayscale = Y.
Parity unpinned: colour() against dtmm itself.  dtmm (v0.7.1, the pinned
source of conda_recipe/recipe.yaml:28-30) is not in /root/reference nor
importable here; colour() restates its published algorithm (XYZ = spec . cmf,
linear sRGB with the D65 primaries, sRGB transfer function with clipping to
[0, 1]; grayscale takes Y) and is pinned only by the closed forms above.
"""
import numpy as np


def _freqs(N, d):
    """Angular frequencies of the FFT bins of N points spaced d (light_propagator.py:603-605)."""
    i = np.arange(N)
    return ((i + 0.5 * N) % N - 0.5 * N) * 2 * np.pi / (N * d)


def objective_filter(Nx, Ny, dx, dy, wl, n_out, qx, qy, NA, zf):
    """Filter plane of one (wavelength, wavevector): mask * exp(i(sqrt(kn^2-k^2)-kn))**|zf|,
    conjugated for zf < 0, without the 1/N of the inverse FFT."""
    k0 = 2 * np.pi / wl
    kn = 2 * np.pi * n_out / wl
    kx = _freqs(Nx, dx)[None, :] + k0 * qx
    ky = _freqs(Ny, dy)[:, None] + k0 * qy
    k2 = kx ** 2 + ky ** 2
    inside = k2 < kn ** 2
    base = np.zeros((Ny, Nx), dtype=np.complex128)
    base[inside] = np.exp(1j * (np.sqrt(kn ** 2 - k2[inside]) - kn))
    f = (k2 < (k0 * NA) ** 2) * base ** np.abs(zf)
    return np.conj(f) if zf < 0 else f


def focus(vals, dx, dy, wavelengths, n_out, wavevectors, NA, z, zfoc_NA_corr=0.0):
    """focused_vals of OpticalFields.focus_fields(z) with objective aperture NA."""
    Nl, Nq, _, Ny, Nx = vals.shape
    n_out = np.broadcast_to(np.asarray(n_out, dtype=np.float64), (Nl,))
    zf = z + NA ** 2 * zfoc_NA_corr
    out = np.empty_like(vals)
    for l in range(Nl):
        for q in range(Nq):
            f = objective_filter(Nx, Ny, dx, dy, wavelengths[l], n_out[l], wavevectors[q][0], wavevectors[q][1],
                                 NA, zf)
            out[l, q] = np.fft.ifft2(np.fft.fft2(vals[l, q], axes=(-2, -1)) * f[None], axes=(-2, -1))
    return out


def _qint(q, qp):
    return q - np.arctanh(q) - qp * np.log(1 - q * q) / 2


def _ring_start(r):
    return 1 + 3 * r * (r - 1)


def quadrature(Nq, Nr, max_NA_condenser, NA, koehler_1D=False):
    """Per-wavevector factors w with spectrum = sum_q w_q s_q (prefactors folded in)."""
    dq = max_NA_condenser / (Nr - 1) if (Nr > 1 and max_NA_condenser > 0) else 1.0
    rmax = min(Nr - 1, int(np.ceil(NA / dq)))
    w = np.zeros(Nq)
    if NA < 1e-8:
        w[Nr - 1 if koehler_1D else 0] = 1.0
        return w
    if not koehler_1D:
        # exact integral of each hat function of q against g(q) = q / (1 - q^2)
        iw = np.zeros(_ring_start(rmax + 1))
        if rmax == 1:
            iw[0] = (_qint(NA, dq) - _qint(0, dq)) / dq
            iw[1:7] = (_qint(0, 0) - _qint(NA, 0)) / (6 * dq)
        else:
            iw[0] = (_qint(dq, dq) - _qint(0, dq)) / dq
            for r in range(1, rmax - 1):
                iw[_ring_start(r):_ring_start(r + 1)] = (
                    _qint((r + 1) * dq, (r + 1) * dq) + _qint((r - 1) * dq, (r - 1) * dq)
                    - 2 * _qint(r * dq, r * dq)) / (6 * r * dq)
            r = rmax - 1
            iw[_ring_start(r):_ring_start(r + 1)] = (
                _qint(NA, rmax * dq) + _qint((rmax - 2) * dq, (rmax - 2) * dq)
                - 2 * _qint(r * dq, r * dq)) / (6 * r * dq)
            iw[_ring_start(rmax):_ring_start(rmax + 1)] = (_qint(r * dq, r * dq) - _qint(NA, r * dq)) / (6 * rmax * dq)
        w[:iw.size] = iw * (2 / NA ** 2)
    else:
        edge = NA - (rmax - 1) * dq
        iw = np.full(2 * rmax + 1, dq)
        iw[[0, -1]] = edge ** 2 / (2 * dq)
        if rmax == 1:
            iw[1] = edge * (2 - edge / dq)
        else:
            iw[[1, -2]] = edge * (1 - edge / (2 * dq)) + dq / 2
        w[Nr - 1 - rmax:Nr + rmax] = iw / (2 * NA)
    return w


def polariser(angle):
    c, s = np.cos(np.deg2rad(angle)), np.sin(np.deg2rad(angle))
    return np.array([[c * c, c * s], [c * s, s * s]], dtype=np.complex128)


def waveplate(kind, angle, wl):
    c, s = np.cos(np.deg2rad(angle)), np.sin(np.deg2rad(angle))
    if kind == "Quarter-wave":
        return np.exp(-0.25j * np.pi) * np.array([[c * c + 1j * s * s, (1 - 1j) * c * s],
                                                  [(1 - 1j) * c * s, s * s + 1j * c * c]])
    if kind == "Half-wave":
        return np.exp(-0.5j * np.pi) * np.array([[c * c - s * s, 2 * c * s], [2 * c * s, s * s - c * c]])
    if kind == "Tint-sensitive":
        g = np.exp(1j * np.pi * 0.540 / wl)
        gc = np.conj(g)
        return np.array([[gc * c * c + g * s * s, (gc - g) * c * s], [(gc - g) * c * s, gc * s * s + g * c * c]])
    raise ValueError(kind)


def jones(optics, wl):
    """(left, right) with T = left @ S @ right for one wavelength."""
    right = np.eye(2, dtype=np.complex128)
    left = np.eye(2, dtype=np.complex128)
    if optics.get("polariser", True):
        right = polariser(optics.get("polariser_angle", 0))
    if optics.get("lower_waveplate", "No") != "No":
        right = waveplate(optics["lower_waveplate"], optics.get("lower_waveplate_angle", 0), wl) @ right
    if optics.get("upper_waveplate", "No") != "No":
        left = waveplate(optics["upper_waveplate"], optics.get("upper_waveplate_angle", 0), wl) @ left
    if optics.get("analyser", True):
        left = polariser(optics.get("analyser_angle", 90)) @ left
    return left, right


def spectrum(focused, wavelengths, optics, Nr, max_NA_condenser, koehler_1D=False):
    """q-averaged spectrum (Ny, Nx, Nl) of FieldViewer (before the intensity factor)."""
    Nl, Nq, _, Ny, Nx = focused.shape
    w = quadrature(Nq, Nr, max_NA_condenser, optics.get("NA_condenser", 0), koehler_1D)
    out = np.zeros((Ny, Nx, Nl))
    for l in range(Nl):
        left, right = jones(optics, wavelengths[l])
        for q in np.nonzero(w)[0]:
            # sample Jones matrix: columns = input polarisation x, y; rows = Ex, Ey
            S = np.stack([np.stack([focused[l, q, 0], focused[l, q, 2]], -1),
                          np.stack([focused[l, q, 1], focused[l, q, 3]], -1)], -2)
            T = left @ S @ right
            out[:, :, l] += w[q] * 0.5 * np.sum(np.abs(T) ** 2, axis=(-1, -2))
    return out


XYZ_TO_LINEAR_SRGB = np.array([[3.2404542, -1.5371385, -0.4985314],
                               [-0.9692660, 1.8760108, 0.0415560],
                               [0.0556434, -0.2040259, 1.0572252]])


def srgb_gamma(v):
    v = np.asarray(v, dtype=np.float64)
    lo = np.where(v < 0, 0.0, 12.92 * v)
    hi = np.where(v > 1, 1.0, 1.055 * np.abs(v) ** (1 / 2.4) - 0.055)
    return np.where(v > 0.0031308, hi, lo)


def colour(spec, cmf, intensity=1.0, grayscale=False, gamma=True):
    """Image (Ny, Nx, 3) of a spectrum (Ny, Nx, Nl) and a colour matching table (Nl, 3)."""
    xyz = (intensity * spec) @ np.asarray(cmf, dtype=np.float64)
    if grayscale:
        rgb = np.repeat(xyz[..., 1:2], 3, axis=-1)
    else:
        rgb = xyz @ XYZ_TO_LINEAR_SRGB.T
    return srgb_gamma(rgb) if gamma else np.clip(rgb, 0, 1)
