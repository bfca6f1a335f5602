"""The default colour matching table of the micrograph synthesis
(nemaktis_amd.colour), used when FieldViewer gets no cmf.  The reference takes
dtmm's tabulated table (field_viewer.py:94-97); dtmm is not importable here, so
the table is pinned by what fixes it independently of this code:

  * the CIE 1931 2-degree functions against the published CIE table at five
    wavelengths (the analytic fit is within 0.01 of them);
  * the black-body shape against Wien's displacement law;
  * the normalisation: the table sums to the D65 white point, so a flat unit
    spectrum gives sRGB white through the oracle's colour conversion.

Against dtmm itself the table is parity unpinned (fit error of the functions
and the black-body approximation of D65).
"""
import numpy as np
import pytest

import viewer_oracle as vo
from nemaktis_amd import colour

# CIE 1931 2-degree observer, published table values (x-bar, y-bar, z-bar)
CIE1931 = {450: (0.3362, 0.0380, 1.7721), 500: (0.0049, 0.3230, 0.2720), 550: (0.4334, 0.9950, 0.0087),
           600: (1.0622, 0.6310, 0.0008), 650: (0.2835, 0.1070, 0.0000)}


def test_cie_functions_match_the_published_table():
    for lam, ref in CIE1931.items():
        np.testing.assert_allclose(colour.cie1931(lam), ref, atol=0.012, err_msg=str(lam))
    # x-bar(600 nm) / y-bar(555 nm) peaks
    assert abs(colour.cie1931(555.0)[1] - 1.0) < 0.005


def test_black_body_peak_follows_wien():
    x = np.linspace(300.0, 900.0, 60001)
    peak = x[np.argmax(colour.planck(x, 6504.0))]
    assert abs(peak - 2.897771955e-3 / 6504.0 * 1e9) < 0.05


@pytest.mark.parametrize("wl", [np.linspace(0.40, 0.80, 21), np.array([0.6, 0.45, 0.55]), np.array([0.55])])
def test_table_is_normalised_to_d65_white(wl):
    t = colour.default_cmf(wl)
    assert t.shape == (wl.size, 3) and np.all(t >= 0)
    np.testing.assert_allclose(t.sum(axis=0), colour.D65_WHITE, rtol=1e-12)
    img = vo.colour(np.ones((2, 3, wl.size)), t, 1.0, gamma=True)
    np.testing.assert_allclose(img, 1.0, atol=1e-6)


def test_table_follows_the_wavelengths_not_their_order():
    wl = np.array([0.45, 0.5, 0.6, 0.7])
    perm = np.array([2, 0, 3, 1])
    np.testing.assert_array_equal(colour.default_cmf(wl[perm]), colour.default_cmf(wl)[perm])
    # blue light has its weight in z, red light in x
    t = colour.default_cmf(wl)
    assert np.argmax(t[0]) == 2 and np.argmax(t[-1]) == 0


def test_invalid_wavelengths_are_refused():
    with pytest.raises(ValueError):
        colour.default_cmf([])
    with pytest.raises(ValueError):
        colour.default_cmf([0.5, -0.1])
    with pytest.raises(ValueError):
        colour.default_cmf([5.0, 6.0])  # infrared: no visible weight
