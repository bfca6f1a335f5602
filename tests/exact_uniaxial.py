"""Exact propagation of transverse fields through a uniform anisotropic medium,
for checking the beam propagation method against Maxwell's equations.

Test infrastructure: an independent reference, derived here from the wave
equation (not a restatement of the propagation code).  In units of the vacuum
wavenumber k0, a plane wave E exp(i k0 (px x + py y + s z)) in a medium of
relative permittivity eps solves

    M(s) E = 0,   M(s) = eps - |k|^2 I + k k^T,   k = (px, py, s),

so the longitudinal wavenumbers s are the four roots of the quartic det M(s)
(its coefficients are found by evaluating the determinant at five points).
The two forward waves (largest Re s + Im s: propagating towards +z, or decaying
there) carry each transverse Fourier mode; their polarisations are null
vectors of M(s) (smallest singular vectors), their transverse parts form the
columns of T, and a distance z multiplies the mode by
T diag(exp(i k0 s_j z)) T^-1.

The same check is the reference's own accuracy benchmark
(HighLevelPythonInterface/examples/benchmarks/3D/01_run_bpm_benchmarks.py:112-196,
doc/source/intro/benchmarks.rst:35-79): a Gaussian beam through a uniform
tilted uniaxial crystal, with the error of each paraxial Fourier mode measured
against the exact propagation of the march's own entrance plane.
"""
import numpy as np


def uniaxial_eps(no, ne, theta, phi=0.0):
    """Relative permittivity of a uniaxial medium whose optical axis has polar angle theta and azimuth phi."""
    n = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    return no ** 2 * np.eye(3) + (ne ** 2 - no ** 2) * np.outer(n, n)


def _wave_matrix(eps, px, py, s):
    """M(s) for arrays px, py, s of one shape: (..., 3, 3) complex."""
    k = np.stack([np.broadcast_to(px, s.shape), np.broadcast_to(py, s.shape), s], axis=-1).astype(complex)
    k2 = np.sum(k * k, axis=-1)
    return eps[None, :, :] * np.ones(s.shape + (1, 1)) - k2[..., None, None] * np.eye(3) + k[..., :, None] * k[..., None, :]


def forward_modes(eps, px, py):
    """Longitudinal wavenumbers s (..., 2) and transverse polarisations T (..., 2, 2) (columns) of the two
    forward waves of each transverse wavevector (px, py) (1-D arrays of one length)."""
    px = np.asarray(px, dtype=float)
    py = np.asarray(py, dtype=float)
    nodes = np.array([-2.0, -1.0, 0.0, 1.0, 2.0])
    dets = np.stack([np.linalg.det(_wave_matrix(eps, px, py, np.full(px.shape, t, dtype=complex))) for t in nodes],
                    axis=-1)
    V = np.vander(nodes, 5)                           # rows: t^4 .. t^0
    coef = np.linalg.solve(V, dets.T).T               # (..., 5), highest power first
    comp = np.zeros(px.shape + (4, 4), dtype=complex)
    comp[..., 0, :] = -coef[..., 1:] / coef[..., :1]
    comp[..., 1, 0] = comp[..., 2, 1] = comp[..., 3, 2] = 1.0
    roots = np.linalg.eigvals(comp)
    order = np.argsort(-(roots.real + roots.imag), axis=-1)[..., :2]
    s = np.take_along_axis(roots, order, axis=-1)
    # a double root (both waves alike, e.g. an isotropic medium) is found split by about the square root of
    # the rounding error: its two values are averaged, and the two waves span the 2-D null space of M
    deg = np.abs(s[..., 0] - s[..., 1]) <= 1e-6 * np.maximum(1.0, np.abs(s[..., 0]))
    s[deg] = np.mean(s[deg], axis=-1, keepdims=True)
    vh = [np.linalg.svd(_wave_matrix(eps, px, py, s[..., j]))[2] for j in range(2)]
    v0 = np.conj(vh[0][..., 2, :])
    v1 = np.where(deg[..., None], np.conj(vh[0][..., 1, :]), np.conj(vh[1][..., 2, :]))
    T = np.stack([v0[..., :2], v1[..., :2]], axis=-1)
    T = T / np.linalg.norm(T, axis=-2, keepdims=True)
    return s, T


def propagate(E0, dx, dy, wavelength, eps, z, p_max):
    """Fourier modes of the transverse field E0 (Ny, Nx, 2) on a periodic mesh, propagated by z through the
    uniform medium, for the modes with |p| <= p_max (p = transverse wavevector / k0); returns
    (Ehat(z), Ehat(0), mask) with Ehat of shape (Ny, Nx, 2) (numpy FFT convention) and zero outside the mask."""
    k0 = 2 * np.pi / wavelength
    Ny, Nx = E0.shape[:2]
    px = np.broadcast_to(2 * np.pi * np.fft.fftfreq(Nx, dx)[None, :] / k0, (Ny, Nx))
    py = np.broadcast_to(2 * np.pi * np.fft.fftfreq(Ny, dy)[:, None] / k0, (Ny, Nx))
    mask = px ** 2 + py ** 2 <= p_max ** 2
    F0 = np.fft.fft2(E0, axes=(0, 1))
    s, T = forward_modes(eps, px[mask], py[mask])
    c = np.linalg.solve(T, F0[mask][..., None])[..., 0]
    Fz = np.zeros_like(F0)
    Fz[mask] = np.einsum("mij,mj->mi", T, c * np.exp(1j * k0 * s * z))
    F0m = np.where(mask[..., None], F0, 0)
    return Fz, F0m, mask


def twisted_nematic_jones(no, ne, wavelength, thickness, twist):
    """Jones matrix, relative to the exp(i k0 no z) carrier, of a planar twisted nematic (director
    (cos qz, sin qz, 0), q = twist / thickness) for normally incident plane waves: in the frame turning with
    the director, E' = R(-qz) E obeys dE'/dz = (i k0 (ne - no) e_x e_x^T - q J) E' with J the rotation
    generator, a constant matrix (the Gooch-Tarry solution)."""
    import scipy.linalg
    k0 = 2 * np.pi / wavelength
    q = twist / thickness
    A = 1j * k0 * (ne - no) * np.array([[1.0, 0.0], [0.0, 0.0]]) - q * np.array([[0.0, -1.0], [1.0, 0.0]])
    R = np.array([[np.cos(twist), -np.sin(twist)], [np.sin(twist), np.cos(twist)]])
    return R @ scipy.linalg.expm(thickness * A)
