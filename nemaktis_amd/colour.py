"""Default colour matching table of the micrograph synthesis.

The reference's ``FieldViewer`` falls back on ``dtmm.color.load_tcmf`` when
no ``cmf`` is given (``HighLevelPythonInterface/nemaktis/field_viewer.py:94-97``):
the CIE 1931 2-degree colour matching functions times the D65 illuminant,
integrated over the wavelength bin of each simulated wavelength, from dtmm's
tabulated data files.  dtmm is a third-party package that is not shipped
here, so this table is built from published closed forms instead:

  * the CIE 1931 2-degree functions by the multi-lobe piecewise-Gaussian fit
    of Wyman, Sloan and Shirley, "Simple Analytic Approximations to the CIE
    XYZ Color Matching Functions", JCGT 2(2), 2013 (their equation 2);
  * the illuminant's spectral shape by Planck's law at D65's correlated
    colour temperature (6504 K);
  * each channel scaled so that the sum over the wavelengths is the D65
    white point (X, Y, Z) = (0.95047, 1, 1.08883): a fully transmitting
    pixel renders as sRGB white, as with dtmm's normalised tables.

Tables therefore differ from dtmm's by the fit error of the functions (a few
percent of their peak) and by the illuminant approximation; a caller who has
dtmm passes ``cmf=dtmm.color.load_tcmf(1000 * wavelengths)`` as before.
"""
import numpy as np

D65_WHITE = np.array([0.95047, 1.0, 1.08883])

# (amplitude, centre nm, sigma below the centre, sigma above the centre)
_LOBES = {
    0: ((1.056, 599.8, 37.9, 31.0), (0.362, 442.0, 16.0, 26.7), (-0.065, 501.1, 20.4, 26.2)),
    1: ((0.821, 568.8, 46.9, 40.5), (0.286, 530.9, 16.3, 31.1)),
    2: ((1.217, 437.0, 11.8, 36.0), (0.681, 459.0, 26.0, 13.8)),
}


def cie1931(wavelength_nm):
    """x-bar, y-bar, z-bar of the CIE 1931 2-degree observer (analytic fit) at
    the given wavelengths in nm; shape (..., 3)."""
    lam = np.asarray(wavelength_nm, dtype=np.float64)
    out = np.zeros(lam.shape + (3,))
    for c, lobes in _LOBES.items():
        for amp, mu, s_lo, s_hi in lobes:
            t = (lam - mu) / np.where(lam < mu, s_lo, s_hi)
            out[..., c] += amp * np.exp(-0.5 * t * t)
    return out


def planck(wavelength_nm, temperature=6504.0):
    """Relative spectral radiance of a black body (arbitrary scale)."""
    lam = np.asarray(wavelength_nm, dtype=np.float64) * 1e-9
    c2 = 1.438776877e-2  # h c / k_B in m K
    return lam ** -5 / np.expm1(c2 / (lam * temperature))


def default_cmf(wavelengths_um, step_nm=0.5):
    """(N, 3) transmission colour matching table for the wavelengths (um):
    XYZ weights of each wavelength's bin (edges half-way to the neighbours,
    the outer bins symmetric), normalised to the D65 white point."""
    wl = np.atleast_1d(np.asarray(wavelengths_um, dtype=np.float64)) * 1e3
    if wl.ndim != 1 or wl.size == 0 or not np.all(np.isfinite(wl)) or np.any(wl <= 0):
        raise ValueError("wavelengths must be a non-empty list of positive values")
    order = np.argsort(wl)
    s = wl[order]
    if s.size == 1:
        table = cie1931(s) * planck(s)[:, None]
    else:
        mid = 0.5 * (s[1:] + s[:-1])
        lo = np.concatenate(([s[0] - (mid[0] - s[0])], mid))
        hi = np.concatenate((mid, [s[-1] + (s[-1] - mid[-1])]))
        table = np.zeros((s.size, 3))
        for i in range(s.size):
            n = max(1, int(np.ceil((hi[i] - lo[i]) / step_nm)))
            x = lo[i] + (np.arange(n) + 0.5) * (hi[i] - lo[i]) / n
            table[i] = (cie1931(x) * planck(x)[:, None]).sum(axis=0) * (hi[i] - lo[i]) / n
    tot = table.sum(axis=0)
    if np.any(tot <= 0):
        raise ValueError("the wavelengths lie outside the visible range of the colour matching functions")
    table *= D65_WHITE / tot
    out = np.empty_like(table)
    out[order] = table
    return out
