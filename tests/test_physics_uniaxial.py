"""The oracle against Maxwell's equations: the reference's own accuracy
benchmark (a Gaussian beam through a uniform tilted uniaxial crystal,
HighLevelPythonInterface/examples/benchmarks/3D/01_run_bpm_benchmarks.py:17-75,
112-196; doc/source/intro/benchmarks.rst:35-79), with the exact solution from
``exact_uniaxial`` (plane-wave eigenmodes of the wave equation, pinned below
against the textbook ordinary / extraordinary indices and the walk-off angle).

The march is compared with the exact propagation of its own entrance plane
(bulk plane 0) to the exit plane, mode by mode in transverse Fourier space
over the paraxial modes |p| <= 0.4 (p = transverse wavevector / k0, the modes
an NA 0.4 objective collects), after removing the exp(i k0 no z) carrier the
march leaves out (01_run_bpm_benchmarks.py:133-137).  Measured (mean paraxial
error as the reference defines it, x- / y-polarised input): N = 21 0.318 /
0.263, N = 41 0.114 / 0.086, N = 81 0.041 / 0.033, N = 161 0.025 / 0.020; the
x-polarised (extraordinary) beam walks off by 1.19 / 1.42 / 1.48 / 1.50 um
against the exact 1.51 um.  On N x 1 x N meshes (one transverse axis) the
errors fall to about 0.01 at N = 161.  A wrong sign or factor in the walk-off
terms moves that centroid by micrometres; a wrong diffraction term breaks the
convergence.  The twisted nematic (uniform beam) converges to the exact Jones
matrix of its continuous twist at second order in dz.
"""
import numpy as np
import pytest

import exact_uniaxial as ex
import oracle
from nemaktis_amd import fields

NO, NE, WL, L, THETA = 1.5, 1.75, 0.6, 10.0, np.pi / 4
P_MAX = 0.4


def _case(N, Ny=None, phi=0.0):
    """N x Ny x N mesh (Ny = N unless given; Ny = 1 is a mesh in the x-z plane) of the 10 um cube, optical
    axis at THETA from z with azimuth phi."""
    Ny = N if Ny is None else Ny
    d = L / (N - 1)
    s = fields.make_settings((N, Ny, N), (d, d, d), [WL], no=str(NO), ne=str(NE), nhost=str(NO), nin=str(NO),
                             beam_profile="GaussianBeam(0.5)", n_woodbury=4, na_objective=0.4)
    lc = np.zeros((N, Ny, N, 3))
    lc[..., 0] = np.sin(THETA) * np.cos(phi)
    lc[..., 1] = np.sin(THETA) * np.sin(phi)
    lc[..., 2] = np.cos(THETA)
    return s, lc, d


def march_errors(bulk, N, d, phi=0.0):
    """(mean paraxial error, relative L2 error, centroid x of the march, exact centroid x) per input
    polarisation, from the march's planes (n_wl, n_q, 2, Nz, 2, Ny, Nx) of a _case mesh."""
    k0 = 2 * np.pi / WL
    eps = ex.uniaxial_eps(NO, NE, THETA, phi)
    Ny = bulk.shape[-2]
    M, My = N - 1, (Ny - 1 if Ny > 1 else 1)  # the march grid: the last column and row repeat the first
    x = (np.arange(M) - M / 2) * d
    out = []
    for pol in (0, 1):
        E0 = np.moveaxis(bulk[0, 0, pol, 0], 0, -1)[:My, :M]
        EL = np.moveaxis(bulk[0, 0, pol, N - 1], 0, -1)[:My, :M]
        Fz, _, mask = ex.propagate(E0, d, d, WL, eps, L, P_MAX)
        Fz = Fz * np.exp(-1j * k0 * NO * L)
        Fb = np.fft.fft2(EL, axes=(0, 1))
        mean_err = np.mean(np.linalg.norm(Fb - Fz, axis=-1)[mask] / (2 * np.linalg.norm(Fz, axis=-1)[mask]))
        rel = np.linalg.norm((Fb - Fz)[mask]) / np.linalg.norm(Fz[mask])

        def centroid(F):
            e = np.fft.ifft2(np.where(mask[..., None], F, 0), axes=(0, 1))
            w = np.sum(np.abs(e) ** 2, axis=-1)
            return float(np.sum(w * x[None, :]) / np.sum(w))

        out.append((mean_err, rel, centroid(Fb), centroid(Fz)))
    return out


def test_exact_modes_untilted_uniaxial_are_the_textbook_indices():
    """Optical axis along z: ordinary s = sqrt(no^2 - p^2) polarised perpendicular to p, extraordinary
    s = no sqrt(1 - p^2 / ne^2) polarised along p."""
    eps = ex.uniaxial_eps(NO, NE, 0.0)
    rng = np.random.default_rng(3)
    p = rng.uniform(-0.6, 0.6, size=(50, 2))
    s, T = ex.forward_modes(eps, p[:, 0], p[:, 1])
    pp = np.hypot(p[:, 0], p[:, 1])
    so, se = np.sqrt(NO ** 2 - pp ** 2), NO * np.sqrt(1 - pp ** 2 / NE ** 2)
    got = np.sort(s.real, axis=-1)
    want = np.sort(np.stack([so, se], axis=-1), axis=-1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10)
    assert np.max(np.abs(s.imag)) < 1e-10
    u = p / pp[:, None]
    for j in range(2):
        along = np.abs(np.einsum("mi,mi->m", T[:, :, j], u))
        is_e = np.abs(s[:, j].real - se) < 1e-9
        np.testing.assert_allclose(along[is_e], 1.0, atol=1e-9)
        np.testing.assert_allclose(along[~is_e], 0.0, atol=1e-9)


def test_exact_modes_tilted_axis_walk_off_angle():
    """Optical axis at theta from z in the x-z plane: at normal incidence the extraordinary index is
    1/sqrt(cos^2/no^2 + sin^2/ne^2), and the slope of its longitudinal wavenumber with px is minus the
    textbook walk-off tan(rho) = (ne^2 - no^2) sin cos / (no^2 sin^2 + ne^2 cos^2)."""
    eps = ex.uniaxial_eps(NO, NE, THETA)
    h = 1e-5
    s, T = ex.forward_modes(eps, np.array([-h, 0.0, h]), np.zeros(3))
    j = int(np.argmax(np.abs(T[1, 0, :])))  # the x-polarised (extraordinary) wave
    n_e = 1 / np.sqrt(np.cos(THETA) ** 2 / NO ** 2 + np.sin(THETA) ** 2 / NE ** 2)
    assert abs(s[1, j].real - n_e) < 1e-10
    assert abs(s[1, 1 - j].real - NO) < 1e-10
    slope = (s[2, j].real - s[0, j].real) / (2 * h)
    tan_rho = (NE ** 2 - NO ** 2) * np.sin(THETA) * np.cos(THETA) / (NO ** 2 * np.sin(THETA) ** 2 +
                                                                    NE ** 2 * np.cos(THETA) ** 2)
    assert abs(-slope - tan_rho) < 1e-6


def test_exact_propagation_isotropic_is_the_angular_spectrum():
    """ne = no: every mode advances by exp(i k0 sqrt(n^2 - p^2) z), whatever its polarisation."""
    eps = ex.uniaxial_eps(NO, NO, THETA)
    rng = np.random.default_rng(5)
    E0 = rng.standard_normal((16, 12, 2)) + 1j * rng.standard_normal((16, 12, 2))
    d, z = 0.3, 2.5
    Fz, F0, mask = ex.propagate(E0, d, d, WL, eps, z, 0.9)
    k0 = 2 * np.pi / WL
    px = 2 * np.pi * np.fft.fftfreq(12, d)[None, :] / k0
    py = 2 * np.pi * np.fft.fftfreq(16, d)[:, None] / k0
    ph = np.exp(1j * k0 * np.sqrt(NO ** 2 - px ** 2 - py ** 2 + 0j) * z)
    want = np.where(mask[..., None], F0 * ph[..., None], 0)
    np.testing.assert_allclose(Fz, want, rtol=0, atol=1e-9 * np.max(np.abs(F0)))


def test_oracle_converges_to_the_exact_tilted_uniaxial_beam():
    errs = {}
    for N in (21, 41, 81):
        s, lc, d = _case(N)
        _, bulk = oracle.propagate_bulk(s, lc)
        errs[N] = march_errors(bulk, N, d)
        print(N, errs[N])
    for pol in (0, 1):
        mean = [errs[N][pol][0] for N in (21, 41, 81)]
        assert mean[0] > mean[1] > mean[2], mean
        assert mean[2] < 0.06 and errs[81][pol][1] < 0.08
    # the extraordinary (x-polarised) beam walks off by about 1.5 um, the ordinary one stays centred
    cx = [errs[N][0][2] for N in (21, 41, 81)]
    exact = errs[81][0][3]
    assert 1.4 < exact < 1.6
    assert abs(cx[2] - exact) < 0.05 and abs(cx[2] - exact) < abs(cx[1] - exact) < abs(cx[0] - exact)
    assert all(abs(errs[N][1][2]) < 0.05 for N in (21, 41, 81))


def test_oracle_converges_on_a_mesh_in_the_xz_plane():
    """The benchmark on N x 1 x N meshes (one transverse axis, 2-D in the reference's benchmark folder), optical
    axis tilted in the mesh plane (walk-off along x) and across it: errors fall to about 1 %."""
    for phi in (0.0, np.pi / 2):
        errs = {}
        for N in (41, 81, 161):
            s, lc, d = _case(N, Ny=1, phi=phi)
            _, bulk = oracle.propagate_bulk(s, lc)
            errs[N] = march_errors(bulk, N, d, phi)
            print(phi, N, errs[N])
        for pol in (0, 1):
            mean = [errs[N][pol][0] for N in (41, 81, 161)]
            assert mean[0] > mean[1] > mean[2] and mean[2] < 0.02, (phi, pol, mean)
        if phi == 0.0:
            assert abs(errs[161][0][2] - errs[161][0][3]) < 0.02 and 1.4 < errs[161][0][3] < 1.6


def tn_case(Nz, no=1.5, ne=1.7, wl=0.6, d=2.0, twist=np.pi / 2):
    """4 x 4 x Nz planar twisted nematic of thickness d (uniform beam: the march reduces to its phase
    operators) and the exact Jones matrix of the continuous twist."""
    dz = d / (Nz - 1)
    s = fields.make_settings((4, 4, Nz), (0.1, 0.1, dz), [wl], no=str(no), ne=str(ne), nhost=str(no), nin=str(no))
    z = np.arange(Nz) * dz
    lc = np.zeros((Nz, 4, 4, 3))
    lc[..., 0] = np.cos(twist * z / d)[:, None, None]
    lc[..., 1] = np.sin(twist * z / d)[:, None, None]
    return s, lc, ex.twisted_nematic_jones(no, ne, wl, d, twist)


def tn_error(bulk, U):
    """Largest relative error of the exit field against U applied to the entrance field (both polarisations),
    and the largest deviation of the exit plane from uniformity."""
    err = dev = 0.0
    for pol in (0, 1):
        E0, EL = bulk[0, 0, pol, 0, :, 1, 1], bulk[0, 0, pol, -1, :, 1, 1]
        want = U @ E0
        err = max(err, np.linalg.norm(EL - want) / np.linalg.norm(want))
        dev = max(dev, np.max(np.abs(bulk[0, 0, pol, -1] - EL[:, None, None])))
    return err, dev


def test_twisted_nematic_converges_to_the_continuum_jones_solution():
    """The twisted-nematic workload's physics (director (cos qz, sin qz, 0), 90 degree twist over 2 um): the
    march converges to the exact Jones matrix of the continuous twist at second order in dz (measured errors
    0.0274, 0.00689, 0.00172, 0.000431 at Nz = 11, 21, 41, 81: ratio 4.00 per halving)."""
    errs = []
    for Nz in (11, 21, 41, 81):
        s, lc, U = tn_case(Nz)
        _, bulk = oracle.propagate_bulk(s, lc)
        err, dev = tn_error(bulk, U)
        assert dev < 1e-13
        errs.append(err)
    print(errs)
    ratios = [a / b for a, b in zip(errs, errs[1:])]
    assert all(3.9 < r < 4.1 for r in ratios), ratios
    assert errs[-1] < 5e-4


@pytest.mark.parametrize("planted", ["walk_off_sign"])
def test_a_planted_walk_off_error_fails_the_benchmark(planted):
    """The same comparison with the optical axis mirrored (theta -> -theta: every walk-off term of the march
    changes sign) must fail the centroid check against the unmirrored exact solution."""
    N = 41
    s, lc, d = _case(N)
    lc[..., 0] = -lc[..., 0]
    _, bulk = oracle.propagate_bulk(s, lc)
    cx, exact = march_errors(bulk, N, d)[0][2:]
    assert abs(cx - exact) > 2.0
