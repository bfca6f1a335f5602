"""Pins of the FP64 CPU oracle against closed forms, textbook special cases,
dense linear algebra and the continuum limit.  The reference repository holds
no golden vectors or known-answer tests for this path, so every check here
comes from what the mathematics fixes; each is chosen so that a plausible
slip in the restatement (a dropped term, a wrong sign or index, a transposed
operand) fails one of them."""
import math

import numpy as np
import pytest
import scipy.linalg

import oracle
from nemaktis_amd import fields

TX, TY, TXY = 0, 1, 2


def _settings(dims, sp=(0.1, 0.1, 0.1), wl=(0.6,), q=((0.0, 0.0),), **kw):
    return fields.make_settings(dims, sp, wl, wavevectors=q, **kw)


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _eps_from_director(n, no, ne):
    de = ne ** 2 - no ** 2
    return no ** 2 * np.eye(3) + de * np.einsum("...i,...j->...ij", n, n)


def _schur_K(eps):
    """Transverse permittivity K = eps_tt - eps_tz eps_zt / eps_zz."""
    return eps[..., :2, :2] - np.einsum("...i,...j->...ij", eps[..., :2, 2], eps[..., 2, :2]) / eps[..., 2, 2][..., None, None]


# --------------------------------------------------------------------------- eps
def test_permittivity_sqrt_and_nref():
    rng = np.random.default_rng(1)
    Nx, Ny, Nz = 7, 6, 3
    n = _unit(rng.standard_normal((Nz, Ny, Nx, 3)))
    mask = rng.uniform(-1, 1, (Nz, Ny, Nx))
    s = _settings((Nx, Ny, Nz), no="1.5", ne="1.7", nhost="1.33")
    ses = oracle.Session(s, n, mask=mask)
    ses.set_wavelength(0)
    tr_sum = 0.0
    for iz in range(Nz):
        e = ses.eps_plane(iz)
        eps = np.zeros((Ny, Nx, 3, 3))
        idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        for c, (i, j) in enumerate(idx):
            eps[..., i, j] = eps[..., j, i] = e[..., c]
        lc = mask[iz] >= 0
        want = np.where(lc[..., None, None], _eps_from_director(n[iz], 1.5, 1.7), 1.33 ** 2 * np.eye(3))
        np.testing.assert_allclose(eps, want, rtol=0, atol=1e-14)
        K = np.stack([np.stack([e[..., 6], e[..., 8]], -1), np.stack([e[..., 8], e[..., 7]], -1)], -2)
        # K is the inverse of the transverse block of eps^-1 (Schur-complement identity)
        np.testing.assert_allclose(K, np.linalg.inv(np.linalg.inv(eps)[..., :2, :2]), atol=1e-12)
        S = np.stack([np.stack([e[..., 9], e[..., 11]], -1), np.stack([e[..., 11], e[..., 10]], -1)], -2)
        np.testing.assert_allclose(S @ S, K, atol=1e-8)
        for iy in range(Ny):
            for ix in range(Nx):
                np.testing.assert_allclose(S[iy, ix], scipy.linalg.sqrtm(K[iy, ix]).real, atol=1e-8)
        tr_sum += (0.5 * (e[..., 9] + e[..., 10])).sum()
    assert abs(ses.nref - tr_sum / (Nx * Ny * Nz)) < 1e-13


# --------------------------------------------------------------------------- phase
@pytest.mark.parametrize("ne_imag", [0.0, 0.02])
def test_phase_operator_is_matrix_exponential(ne_imag):
    rng = np.random.default_rng(2)
    Nx, Ny, Nz = 6, 5, 3
    n = _unit(rng.standard_normal((Nz, Ny, Nx, 3)))
    s = _settings((Nx, Ny, Nz), no="1.5", ne="1.7", ne_imag=str(ne_imag))
    ses = oracle.Session(s, n)
    ses.set_wavelength(0)
    ses.update(1)
    P = ses.phase_coefs()
    k0 = 2 * math.pi / 0.6
    dZ = k0 * 0.1
    alpha = 1 + 1j * ne_imag / (1.7 - 1.5)
    for iy in range(Ny):
        for ix in range(Nx):
            S = []
            for iz in (1, 2):
                K = _schur_K(_eps_from_director(n[iz, iy, ix], 1.5, 1.7))
                S.append(scipy.linalg.sqrtm(K).real)
            d = 0.5 * dZ * (0.5 * (S[0] + S[1]) - 1.5 * np.eye(2))
            want = scipy.linalg.expm(1j * alpha * d)
            got = np.array([[P[0, iy, ix], P[2, iy, ix]], [P[2, iy, ix], P[1, iy, ix]]])
            np.testing.assert_allclose(got, want, atol=1e-7)  # sqrt(K) branch tolerance


# --------------------------------------------------------------------------- solves
@pytest.mark.parametrize("M", [3, 4, 7, 64])
def test_cyclic_line_solve_matches_dense(M):
    rng = np.random.default_rng(M)
    c = lambda: rng.standard_normal(M) + 1j * rng.standard_normal(M)
    a, cc, d = 0.3 * c(), 0.3 * c(), c()
    b = 2.0 + c() * 0.2
    A = np.zeros((M, M), complex)
    for j in range(M):
        A[j, j] += b[j]
        A[j, (j - 1) % M] += a[j]
        A[j, (j + 1) % M] += cc[j]
    x = oracle.cyclic_line_solve(a, b, cc, d)
    np.testing.assert_allclose(x, np.linalg.solve(A, d), rtol=1e-12, atol=1e-12)


def _random_field(rng, Nx, Ny):
    f = rng.standard_normal((2, Ny, Nx)) + 1j * rng.standard_normal((2, Ny, Nx))
    f[:, :, Nx - 1] = f[:, :, 0]
    f[:, Ny - 1, :] = f[:, 0, :]
    return np.ascontiguousarray(f)


@pytest.mark.parametrize("direction,block", [(0, 0), (0, 3), (1, 0), (1, 3)])
def test_block_solve_matches_dense_lines(direction, block):
    rng = np.random.default_rng(5)
    Nx, Ny, Nz = 9, 8, 2
    n = _unit(rng.standard_normal((Nz, Ny, Nx, 3)))
    s = _settings((Nx, Ny, Nz), no="1.5", ne="1.7")
    ses = oracle.Session(s, n)
    ses.set_wavelength(0)
    ses.update(0)
    src = _random_field(rng, Nx, Ny)
    dst = ses.field()
    comp = 0 if block == 0 else 1
    ses.block_solve(direction, block, dst, comp, src, comp)
    t = TX if direction == 0 else TY
    coef = [ses.R(t, block, k) for k in range(3)]
    mu = 1j * (2 * math.pi / 0.6) * 0.1 / 2
    lines = Ny if direction == 0 else Nx
    M = (Nx if direction == 0 else Ny) - 1
    for l in range(lines):
        sl = (lambda j: (l, j)) if direction == 0 else (lambda j: (j, l))
        A = np.eye(M, dtype=complex)
        for j in range(M):
            for k, off in ((0, -1), (1, 0), (2, 1)):
                A[j, (j + off) % M] -= mu * coef[k][sl(j)]
        rhs = np.array([src[comp][sl(j)] for j in range(M)])
        x = np.linalg.solve(A, rhs)
        got = np.array([dst[comp][sl(j)] for j in range(M + 1)])
        np.testing.assert_allclose(got[:M], x, rtol=1e-11, atol=1e-12)
        assert got[M] == got[0]


# --------------------------------------------------------------------------- ADI
def _adi_symbol(theta_x, theta_y, dX, dY, dZ, eps, nw):
    """2x2 amplification matrix of one ADI step for a Fourier mode in a
    homogeneous anisotropic medium, derived from the operator symbols."""
    ezz = eps[2, 2]
    e = 1.0 / ezz
    K = _schur_K(eps)
    S = scipy.linalg.sqrtm(K).real
    Si = np.linalg.inv(S)
    vx, vy = eps[0, 2] / ezz, eps[1, 2] / ezz
    px = 1j * np.sin(theta_x) / dX
    py = 1j * np.sin(theta_y) / dY
    lx = (2 * np.cos(theta_x) - 2) / dX ** 2
    ly = (2 * np.cos(theta_y) - 2) / dY ** 2
    mx = -np.sin(theta_x) * np.sin(theta_y) / (dX * dY)
    kxx, kyy, kxy = K[0, 0], K[1, 1], K[0, 1]
    D = {TX: np.array([[e * lx * kxx, e * lx * kxy], [0, lx]], complex),
         TY: np.array([[ly, 0], [e * ly * kxy, e * ly * kyy]], complex),
         TXY: np.array([[e * mx * kxy, -mx + e * mx * kyy], [-mx + e * mx * kxx, e * mx * kxy]], complex)}
    W = {TX: 0.5 * np.array([[2 * vx * px, vy * px], [vy * px, 0]]),
         TY: 0.5 * np.array([[0, vx * py], [vx * py, 2 * vy * py]]),
         TXY: np.zeros((2, 2))}
    R = {t: 1j * W[t] + 0.25 * (D[t] @ Si + Si @ D[t]) for t in (TX, TY, TXY)}
    mu = 1j * dZ / 2
    I = np.eye(2)
    Ax = R[TX] + np.array([R[TXY][0], [0, 0]])
    Ay = R[TY] + np.array([[0, 0], R[TXY][1]])

    def back(A, T):
        O = A - T
        M = np.linalg.inv(I - mu * T)
        acc, term = M.copy(), M.copy()
        for _ in range(nw):
            term = M @ (mu * O) @ term
            acc = acc + term
        return acc
    Tx = np.diag([R[TX][0, 0], R[TX][1, 1]])
    Ty = np.diag([R[TY][0, 0], R[TY][1, 1]])
    return back(Ay, Ty) @ (I + mu * Ax) @ back(Ax, Tx) @ (I + mu * Ay)


@pytest.mark.parametrize("nw", [0, 2, 4])
def test_adi_step_fourier_symbol_homogeneous_anisotropic(nw):
    rng = np.random.default_rng(7)
    Nx, Ny, Nz = 17, 13, 2
    th, ph = 0.7, 0.45
    d = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    n = np.broadcast_to(d, (Nz, Ny, Nx, 3)).copy()
    s = _settings((Nx, Ny, Nz), sp=(0.1, 0.12, 0.1), no="1.5", ne="1.7", n_woodbury=nw)
    ses = oracle.Session(s, n)
    ses.set_wavelength(0)
    ses.update(0)
    f = _random_field(rng, Nx, Ny)
    g = f.copy()
    ses.apply_primary(g)
    Mx, My = Nx - 1, Ny - 1
    fh = np.fft.fft2(f[:, :My, :Mx])
    k0 = 2 * math.pi / 0.6
    eps = _eps_from_director(d, 1.5, 1.7)
    out = np.empty_like(fh)
    for my in range(My):
        for mxx in range(Mx):
            U = _adi_symbol(2 * math.pi * mxx / Mx, 2 * math.pi * my / My, k0 * 0.1, k0 * 0.12, k0 * 0.1, eps, nw)
            out[:, my, mxx] = U @ fh[:, my, mxx]
    want = np.fft.ifft2(out)
    err = np.linalg.norm(g[:, :My, :Mx] - want) / np.linalg.norm(want)
    assert err < 1e-12, err


def test_primary_operator_continuum_limit():
    """The assembled R = iW + (D S^-1 + S^-1 D)/4 of a smooth inhomogeneous
    medium converges at second order to the continuous operator
    (D f)_a = (D0 f)_a + d_a(ezz^-1 div(K f)), (W f)_a = (v_a div f + d_a(v.f))/2."""
    errs = []
    for M in (24, 48):
        Nx = Ny = M + 1
        L = 3.2
        h = L / M
        x = np.arange(Nx) * h
        X, Y = np.meshgrid(x, x, indexing="xy")
        th = 0.6 + 0.3 * np.sin(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / L)
        ph = 0.4 + 0.5 * np.cos(2 * np.pi * X / L + 0.3) + 0.2 * np.sin(2 * np.pi * Y / L)
        n2 = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
        n = np.stack([n2, n2])
        s = _settings((Nx, Ny, 2), sp=(h, h, 0.1), no="1.5", ne="1.7")
        ses = oracle.Session(s, n)
        ses.set_wavelength(0)
        ses.update(0)
        f = np.stack([np.exp(1j * (np.sin(2 * np.pi * X / L) + 0.5 * np.cos(2 * np.pi * Y / L))),
                      (1 + 0.3 * np.sin(2 * np.pi * Y / L)) * np.exp(-1j * np.cos(2 * np.pi * X / L))])
        f = np.ascontiguousarray(f)
        Rf = ses.field()
        first = True
        for t in (TX, TY, TXY):
            for b in range(4):
                ses.block_vmult(t, b, Rf, b // 2, f, b % 2, add=True)
        # continuum operator with spectral derivatives on the periodic interior grid
        k0 = 2 * math.pi / 0.6
        kk = 2 * np.pi * np.fft.fftfreq(M, d=h * k0)

        def dx(a):
            return np.fft.ifft(1j * kk[None, :] * np.fft.fft(a, axis=1), axis=1)

        def dy(a):
            return np.fft.ifft(1j * kk[:, None] * np.fft.fft(a, axis=0), axis=0)
        ni = n2[:M, :M]
        eps = _eps_from_director(ni, 1.5, 1.7)
        K = _schur_K(eps)
        S = np.zeros_like(K)
        for i in range(M):
            for j in range(M):
                S[i, j] = scipy.linalg.sqrtm(K[i, j]).real
        Si = np.linalg.inv(S)
        e = 1.0 / eps[..., 2, 2]
        vx, vy = eps[..., 0, 2] * e, eps[..., 1, 2] * e
        fi = f[:, :M, :M]
        mv = lambda A, v: np.stack([A[..., 0, 0] * v[0] + A[..., 0, 1] * v[1], A[..., 1, 0] * v[0] + A[..., 1, 1] * v[1]])

        def Dop(g):
            Kg = mv(K, g)
            div = dx(Kg[0]) + dy(Kg[1])
            return np.stack([dy(dy(g[0])) - dx(dy(g[1])) + dx(e * div),
                             -dx(dy(g[0])) + dx(dx(g[1])) + dy(e * div)])
        divf = dx(fi[0]) + dy(fi[1])
        vf = vx * fi[0] + vy * fi[1]
        Wf = 0.5 * np.stack([vx * divf + dx(vf), vy * divf + dy(vf)])
        want = 1j * Wf + 0.25 * (Dop(mv(Si, fi)) + mv(Si, Dop(fi)))
        errs.append(np.linalg.norm(Rf[:, :M, :M] - want) / np.linalg.norm(want))
    assert errs[1] < 0.03, errs
    assert errs[0] / errs[1] > 3.2, errs  # second-order convergence


# --------------------------------------------------------------------------- full march
def test_xy_invariant_march_is_product_of_2x2_matrices():
    Nx, Ny, Nz = 8, 8, 6
    n = fields.twisted_nematic(Nx, Ny, Nz)
    s = _settings((Nx, Ny, Nz), no="1.5", ne="1.7", ne_imag="0.01", nin="1.2", nout="1.3", na_objective=0.5)
    E = oracle.propagate(s, n)
    k0 = 2 * math.pi / 0.6
    dZ = k0 * 0.1
    alpha = 1 + 1j * 0.01 / 0.2
    S = [scipy.linalg.sqrtm(_schur_K(_eps_from_director(n[iz, 0, 0], 1.5, 1.7))).real for iz in range(Nz)]
    I = np.eye(2)
    T = 2 * 1.2 * np.linalg.inv(1.2 * I + S[0])
    for iz in range(Nz - 1):
        P = scipy.linalg.expm(1j * alpha * 0.5 * dZ * (0.5 * (S[iz] + S[iz + 1]) - 1.5 * I))
        T = P @ P @ T
    T = 2 * S[-1] @ np.linalg.inv(1.3 * I + S[-1]) @ T
    for pol in range(2):
        e_in = np.array([math.cos(0.5 * math.pi * pol), math.sin(0.5 * math.pi * pol)])
        want = T @ e_in
        got = E[0, 0, 2 * pol:2 * pol + 2]
        np.testing.assert_allclose(got, np.broadcast_to(want[:, None, None], got.shape), rtol=0, atol=1e-12)


def test_uniaxial_slab_retardation_and_absorption():
    """A uniform uniaxial slab with the optic axis along x at normal incidence
    (the textbook wave plate): the x-polarised (extraordinary) and y-polarised
    (ordinary) waves keep their polarisation, each passes both interfaces
    with its own normal-incidence transmissions, the extraordinary wave gains
    the retardation k0 (ne - no) L over the ordinary one (whose phase the march
    takes as its reference) and is damped by exp(-k0 ne_imag L): ne_imag is
    the imaginary part of the extraordinary index (Beer-Lambert)."""
    Nx, Ny, Nz = 6, 5, 11
    no, ne, ki, nin, nout = 1.5, 1.7, 0.01, 1.2, 1.3
    s = _settings((Nx, Ny, Nz), no=str(no), ne=str(ne), ne_imag=str(ki), nin=str(nin), nout=str(nout),
                  na_objective=0.5)
    n = np.zeros((Nz, Ny, Nx, 3)); n[..., 0] = 1
    E = oracle.propagate(s, n)
    k0, L = 2 * math.pi / 0.6, 0.1 * (Nz - 1)
    t = lambda nj: 2 * nin / (nin + nj) * 2 * nj / (nj + nout)
    np.testing.assert_allclose(E[0, 0, 0], t(ne) * np.exp(1j * k0 * (ne - no) * L - k0 * ki * L), rtol=0, atol=1e-12)
    np.testing.assert_allclose(E[0, 0, 3], t(no), rtol=0, atol=1e-12)
    assert np.abs(E[0, 0, 1]).max() < 1e-14 and np.abs(E[0, 0, 2]).max() < 1e-14


def test_isotropic_slab_textbook_transmission():
    Nx, Ny, Nz = 6, 6, 5
    n = np.zeros((Nz, Ny, Nx, 3)); n[..., 2] = 1
    mask = -np.ones((Nz, Ny, Nx))
    nh, nin, nout = 1.6, 1.0, 1.3
    s = _settings((Nx, Ny, Nz), no="1.5", ne="1.7", nhost=str(nh), nin=str(nin), nout=str(nout))
    E = oracle.propagate(s, n, mask=mask)
    dZ = 2 * math.pi / 0.6 * 0.1
    t = 2 * nin / (nin + nh) * 2 * nh / (nh + nout) * np.exp(1j * dZ * (nh - 1.5) * (Nz - 1))
    for pol in range(2):
        e_in = np.array([math.cos(0.5 * math.pi * pol), math.sin(0.5 * math.pi * pol)])
        got = E[0, 0, 2 * pol:2 * pol + 2]
        np.testing.assert_allclose(got, np.broadcast_to((t * e_in)[:, None, None], got.shape), atol=1e-13)


# --------------------------------------------------------------------------- shift
def test_shift_translates_smooth_profile():
    Nx, Ny, Nz = 65, 49, 2
    dx, dy, dz = 0.1, 0.13, 0.2
    q = (0.18, -0.12)
    n = np.zeros((Nz, Ny, Nx, 3)); n[..., 2] = 1
    mask = -np.ones((Nz, Ny, Nx))
    s = _settings((Nx, Ny, Nz), sp=(dx, dy, dz), q=(q,), no="1.5", ne="1.7", nhost="1.5")
    ses = oracle.Session(s, n, mask=mask)
    ses.set_wavelength(0)
    nref = ses.nref
    assert abs(nref - 1.5) < 1e-14
    Mx, My = Nx - 1, Ny - 1
    x = np.arange(Nx) * dx
    y = np.arange(Ny) * dy
    X, Y = np.meshgrid(x, y, indexing="xy")
    Lx, Ly = Mx * dx, My * dy
    prof = lambda X, Y: np.exp(np.cos(2 * np.pi * X / Lx) + np.sin(2 * np.pi * Y / Ly))
    f = np.ascontiguousarray(np.stack([prof(X, Y), 2 * prof(X, Y)]).astype(complex))
    ses.apply_shift(0, f)
    # translation of the pattern by +dz q/(2 nref) along x; along y the reference
    # converts with deltaX, i.e. dz qy dy / (2 nref dx)
    sx = dz * q[0] / (2 * nref)
    sy = dz * q[1] * dy / (2 * nref * dx)
    k0 = 2 * math.pi / 0.6
    ph = np.exp(-1j * k0 * dz * (q[0] ** 2 + q[1] ** 2) / (4 * nref))
    want = ph * prof(X - sx, Y - sy)
    np.testing.assert_allclose(f[0][:My, :Mx], want[:My, :Mx], rtol=2e-3)
    np.testing.assert_allclose(f[1], 2 * f[0], rtol=1e-15)


def test_shift_weights_partition_of_unity_and_linear_precision():
    Nx, Ny, Nz = 40, 4, 2
    q = (0.37, 0.0)
    n = np.zeros((Nz, Ny, Nx, 3)); n[..., 2] = 1
    s = _settings((Nx, Ny, Nz), q=(q,), no="1.5", ne="1.7", nhost="1.5")
    ses = oracle.Session(s, n, mask=-np.ones((Nz, Ny, Nx)))
    ses.set_wavelength(0)
    ramp = np.arange(Nx, dtype=float)
    f = np.ascontiguousarray(np.stack([np.broadcast_to(ramp, (Ny, Nx)), np.ones((Ny, Nx))]).astype(complex))
    ses.apply_shift(0, f)
    k0 = 2 * math.pi / 0.6
    ph = np.exp(-1j * k0 * 0.1 * q[0] ** 2 / (4 * 1.5))
    cells = -k0 * 0.1 * q[0] / (2 * 1.5 * k0 * 0.1)
    np.testing.assert_allclose(f[1] / ph, 1.0, atol=1e-14)
    inner = slice(4, Nx - 5)
    np.testing.assert_allclose((f[0] / ph)[0, inner].real, ramp[inner] + cells, atol=1e-12)


# --------------------------------------------------------------------------- filter
@pytest.mark.parametrize("Nx,Ny", [(16, 12), (15, 10)])
def test_oracle_fft_matches_numpy(Nx, Ny):
    rng = np.random.default_rng(9)
    a = rng.standard_normal((3, Ny, Nx)) + 1j * rng.standard_normal((3, Ny, Nx))
    np.testing.assert_allclose(oracle.fft2(a, -1), np.fft.fft2(a), rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(oracle.fft2(a, +1), np.fft.ifft2(a) * Nx * Ny, rtol=1e-12, atol=1e-11)


def test_filter_without_layers_is_na_lowpass():
    rng = np.random.default_rng(11)
    Nx, Ny = 24, 20
    q = ((0.0, 0.0), (0.1, -0.05))
    s = _settings((Nx, Ny, 2), sp=(0.2, 0.25, 0.1), q=q, no="1.5", ne="1.7", na_objective=0.55)
    ses = oracle.Session(s, np.zeros((2, Ny, Nx, 3)) + [0, 0, 1])
    E = rng.standard_normal((2, 4, Ny, Nx)) + 1j * rng.standard_normal((2, 4, Ny, Nx))
    got = np.ascontiguousarray(E.copy())
    ses.filter_wavelength(0, got)
    k0 = 2 * math.pi / 0.6
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=0.2) / k0
    ky = 2 * np.pi * np.fft.fftfreq(Ny, d=0.25) / k0
    for iq, (qx, qy) in enumerate(q):
        m = np.hypot(kx[None, :] + qx, ky[:, None] + qy) < 0.55
        want = np.fft.ifft2(np.fft.fft2(E[iq]) * m)
        np.testing.assert_allclose(got[iq], want, atol=1e-12)


def test_filter_single_layer_normal_incidence_fresnel():
    n1, n2, e_thick = 1.45, 1.2, 3.0
    Nx, Ny = 8, 8
    s = _settings((Nx, Ny, 2), no="1.5", ne="1.7", nout=str(n2), na_objective=0.9,
                  iso_thicknesses=[e_thick], iso_indices=[n1])
    ses = oracle.Session(s, np.zeros((2, Ny, Nx, 3)) + [0, 0, 1])
    E = np.ones((1, 4, Ny, Nx), complex) * np.array([1.0, 2.0, -1.0, 0.5])[None, :, None, None]
    got = np.ascontiguousarray(E.copy())
    ses.filter_wavelength(0, got)
    np.testing.assert_allclose(got, E * 2 * n1 / (n1 + n2), atol=1e-13)


def test_filter_index_matched_layer_is_angular_spectrum_propagator():
    rng = np.random.default_rng(13)
    n, zf = 1.3, 2.5
    Nx, Ny = 16, 16
    s = _settings((Nx, Ny, 2), sp=(0.3, 0.3, 0.1), no="1.5", ne="1.7", nout=str(n), na_objective=0.8,
                  iso_thicknesses=[1.7], iso_indices=[n], z_shift=zf)
    ses = oracle.Session(s, np.zeros((2, Ny, Nx, 3)) + [0, 0, 1])
    E = rng.standard_normal((1, 4, Ny, Nx)) + 1j * rng.standard_normal((1, 4, Ny, Nx))
    got = np.ascontiguousarray(E.copy())
    ses.filter_wavelength(0, got)
    k0 = 2 * math.pi / 0.6
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=0.3) / k0
    qq = kx[None, :] ** 2 + kx[:, None] ** 2
    # evanescent q (beyond n) lie outside the aperture, where H is 0 anyway
    kz = np.sqrt(np.maximum(n * n - qq, 0.0))
    H = np.where(np.sqrt(qq) < 0.8, np.exp(1j * (kz - n) * k0 * zf), 0)
    want = np.fft.ifft2(np.fft.fft2(E[0]) * H)
    np.testing.assert_allclose(got[0], want, atol=1e-11)


def test_focus_correction_takes_the_exit_index_at_0_6_um():
    """The focus correction subtracts sum(e_p * nout(0.6 um) / n_p) over the
    isotropic layers, at every wavelength (ScreenOutput.cpp:34-38).  A layer
    index-matched to the exit medium at the propagated wavelength (0.5 um) is
    more of that medium, so the filter must equal the angular-spectrum
    propagator over z_foc + e - e * nout(0.6) / nout(0.5); with nout taken at
    the propagated wavelength the layer would cancel exactly (over z_foc)."""
    rng = np.random.default_rng(17)
    nout = "1.3+0.01/lambda^2"
    n5, n6 = 1.3 + 0.01 / 0.25, 1.3 + 0.01 / 0.36
    zf, e = 2.5, 1.7
    Nx, Ny = 16, 16
    s = _settings((Nx, Ny, 2), sp=(0.3, 0.3, 0.1), wl=(0.5,), no="1.5", ne="1.7", nout=nout, na_objective=0.8,
                  iso_thicknesses=[e], iso_indices=[n5], z_shift=zf)
    ses = oracle.Session(s, np.zeros((2, Ny, Nx, 3)) + [0, 0, 1])
    E = rng.standard_normal((1, 4, Ny, Nx)) + 1j * rng.standard_normal((1, 4, Ny, Nx))
    got = np.ascontiguousarray(E.copy())
    ses.filter_wavelength(0, got)
    k0 = 2 * math.pi / 0.5
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=0.3) / k0
    qq = kx[None, :] ** 2 + kx[:, None] ** 2
    kz = np.sqrt(np.maximum(n5 * n5 - qq, 0.0))

    def propagated(z):
        H = np.where(np.sqrt(qq) < 0.8, np.exp(1j * (kz - n5) * k0 * z), 0)
        return np.fft.ifft2(np.fft.fft2(E[0]) * H)

    np.testing.assert_allclose(got[0], propagated(zf + e - e * n6 / n5), atol=1e-10)
    assert np.abs(got[0] - propagated(zf)).max() > 1e-3


# --------------------------------------------------------------------------- Q-tensor input
def test_qtensor_of_a_director_reproduces_director_mode():
    """PermittivityTensorField.cpp:41-65 (Q-tensor branch) against :18-40 (director
    branch): the caller's Q of a uniaxial director of order 1 is (3 n n^T - I)/2
    (HighLevelPythonInterface/nemaktis/lc_material.py:667-681), for which
    (no^2 + de/3) I + (2 de/3) Q = no^2 I + de n n^T exactly.  The two input modes
    must then give the same permittivity and the same output fields, so a wrong
    isotropic part or anisotropy factor of the Q branch fails here."""
    dims = (14, 12, 5)
    n = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=21, corr_length=0.3)
    Q = 0.5 * (3 * np.einsum("...i,...j->...ij", n, n) - np.eye(3))
    q6 = np.stack([Q[..., 0, 0], Q[..., 1, 1], Q[..., 2, 2], Q[..., 0, 1], Q[..., 0, 2], Q[..., 1, 2]], -1)
    kw = dict(wl=(0.55, 0.63), q=((0.0, 0.0), (0.1, -0.05)), no="1.5", ne="1.7", nhost="1.4", na_objective=0.6)
    s_dir = _settings(dims, **kw)
    s_q = _settings(dims, field_type="Q-tensor", **kw)
    mask = np.where(np.arange(dims[0])[None, None, :] < 10, 1.0, -1.0) * np.ones((dims[2], dims[1], 1))
    a = oracle.Session(s_dir, n, mask=mask)
    b = oracle.Session(s_q, q6, mask=mask)
    for w in range(2):
        a.set_wavelength(w)
        b.set_wavelength(w)
        assert abs(a.nref - b.nref) < 1e-14
        for iz in range(dims[2]):
            np.testing.assert_allclose(b.eps_plane(iz), a.eps_plane(iz), rtol=0, atol=2e-14)
    e_dir = oracle.propagate(s_dir, n, mask=mask)
    e_q = oracle.propagate(s_q, q6, mask=mask)
    assert np.linalg.norm(e_q - e_dir) / np.linalg.norm(e_dir) < 1e-13
    # a Q of lower order (S = 0.6) is a different medium: the check is not vacuous
    e_s = oracle.propagate(s_q, 0.6 * q6, mask=mask)
    assert np.linalg.norm(e_s - e_dir) / np.linalg.norm(e_dir) > 1e-3


# --------------------------------------------------------------------------- mask-formula weights
def _crossing_weights(z_root, dims, sp):
    """Closed-form weights: the crossing z* of each slab it falls in, (z* - z_iz)/dz."""
    Nx, Ny, Nz = dims
    dz = sp[2]
    zi = (np.arange(Nz - 1) - 0.5 * (Nz - 1)) * dz
    w = np.full((Nz - 1, Ny, Nx, 2), 0.5)
    for iz in range(Nz - 1):
        inside = (z_root >= zi[iz]) & (z_root <= zi[iz] + dz)
        w0 = (z_root - zi[iz]) / dz
        w[iz, ..., 0] = np.where(inside, w0, 0.5)
        w[iz, ..., 1] = np.where(inside, 1 - w0, 0.5)
    return w


def test_oracle_mask_weights_planar_interface_closed_form():
    """A planar interface z = z0 inside slab k: every column's weights of slab k
    are (z0 - z_k)/dz up to the bisection bracket (1e-3 dz), all others 0.5."""
    import mask_weights
    dims, sp = (5, 4, 7), (0.2, 0.2, 0.25)
    z = (np.arange(dims[2]) - 0.5 * (dims[2] - 1)) * sp[2]
    for z0 in (0.11, -0.37, 0.6):
        mask = (z0 - z)[:, None, None] * np.ones((1, dims[1], dims[0]))
        w, bad = mask_weights.interface_weights(lambda x, y, zz: z0 - zz + 0 * x, dims, sp, mask)
        assert not bad
        want = _crossing_weights(np.full((dims[1], dims[0]), z0), dims, sp)
        np.testing.assert_allclose(w, want, rtol=0, atol=1e-3)
        k = int(np.floor((z0 - z[0]) / sp[2]))
        assert np.all(np.abs(w[k, ..., 0] - (z0 - z[k]) / sp[2]) < 1e-3)
        assert np.all(np.delete(w, k, axis=0) == 0.5)


def test_oracle_mask_weights_spherical_interface_closed_form():
    """A droplet R^2 - r^2: crossings at z* = +-sqrt(R^2 - x^2 - y^2)."""
    import mask_weights
    dims, sp, R = (11, 10, 12), (0.3, 0.3, 0.3), 1.4
    x, y, z = fields.mesh_coords(*dims, *sp)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    mask = R * R - X * X - Y * Y - Z * Z
    w, bad = mask_weights.interface_weights(lambda a, b, c: R * R - a * a - b * b - c * c, dims, sp, mask)
    assert not bad
    rho2 = R * R - x[None, :] ** 2 - y[:, None] ** 2
    with np.errstate(invalid="ignore"):
        root = np.sqrt(rho2)
    want = np.full_like(w, 0.5)
    for sgn in (1.0, -1.0):
        r = np.where(rho2 > 0, sgn * root, np.nan)
        cw = _crossing_weights(r, dims, sp)
        hit = cw[..., 0] != 0.5
        want[hit] = cw[hit]
    np.testing.assert_allclose(w, want, rtol=0, atol=1e-3)
    assert np.sum(w[..., 0] != 0.5) > 20


def test_oracle_mask_weights_mismatch_falls_back():
    """A formula whose crossings disagree with the mask: every weight 0.5
    (UtilitiesLib/VectorField.cpp:107-111)."""
    import mask_weights
    dims, sp = (6, 5, 6), (0.2, 0.2, 0.2)
    w, bad = mask_weights.interface_weights(lambda x, y, z: 0.05 - z + 0 * x, dims, sp, np.ones(dims[::-1]))
    assert bad and np.all(w == 0.5)


def test_oracle_single_plane_is_the_two_fresnel_interfaces():
    """A mesh of one plane (Nz = 1): no slab is marched (BPMIteration.cpp:127
    loops Nz-1 times), so a uniform beam through an isotropic host medium of
    index n comes out multiplied by the textbook transmissions
    2 nin / (nin + n) (FresnelOperator.cpp:16-38) and 2 n / (n + nout)
    (:45-69); the screen filter passes the uniform field (q = 0 inside the
    objective NA, no focus shift)."""
    dims, sp = (6, 5, 1), (0.1, 0.1, 0.1)
    nin, n, nout = 1.2, 1.45, 1.33
    s = fields.make_settings(dims, sp, [0.6], no="1.5", ne="1.7", nhost=str(n), nin=str(nin), nout=str(nout),
                             na_objective=0.5)
    lc = fields.random_smooth(*dims, sp, seed=3, corr_length=0.2)
    E = oracle.propagate(s, lc, mask=-np.ones(dims[::-1]))
    t = 4 * nin * n / ((nin + n) * (n + nout))
    np.testing.assert_allclose(E[0, 0, 0], t, rtol=1e-13)        # x input -> Ex
    np.testing.assert_allclose(E[0, 0, 3], t, rtol=1e-13)        # y input -> Ey
    assert np.abs(E[0, 0, 1]).max() < 1e-15 and np.abs(E[0, 0, 2]).max() < 1e-13


def test_oracle_single_plane_anisotropic_interfaces_are_the_eigenmode_transmissions():
    """A mesh of one plane of uniaxial LC with an in-plane director at angle
    phi: S = sqrt(K) has the eigen-polarisations along and across the
    director with indices ne and no, and each passes both interfaces with the
    textbook normal-incidence transmissions, 2 nin / (nin + n_j) then
    2 n_j / (n_j + nout) (FresnelOperator.cpp:16-38, 45-69): the output is
    R diag(T_e, T_o) R^T applied to the two input polarisations.  A sign slip
    on S_xy or a transposed adjugate fails it for phi off the axes."""
    dims, sp = (6, 5, 1), (0.1, 0.1, 0.1)
    nin, nout, no, ne, phi = 1.2, 1.33, 1.5, 1.7, 0.7
    s = fields.make_settings(dims, sp, [0.6], no=str(no), ne=str(ne), nin=str(nin), nout=str(nout), na_objective=0.5)
    lc = np.zeros(dims[::-1] + (3,)) + [math.cos(phi), math.sin(phi), 0.0]
    E = oracle.propagate(s, lc)
    R = np.array([[math.cos(phi), -math.sin(phi)], [math.sin(phi), math.cos(phi)]])
    T = R @ np.diag([4 * nin * n / ((nin + n) * (n + nout)) for n in (ne, no)]) @ R.T
    for pol in (0, 1):
        for comp in (0, 1):
            np.testing.assert_allclose(E[0, 0, 2 * pol + comp], T[comp, pol], rtol=0, atol=1e-13)


# --------------------------------------------------------------------------- seam rule
def _tilted(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


@pytest.mark.parametrize("axis", ["x", "y"])
def test_duplicate_plane_enters_the_march_only_through_eps_zz(axis):
    """The march grid excludes the duplicate column (x) or row (y) of the mesh,
    but the D1 seam rule reads 1/eps_zz there for the +x / +y neighbour of the
    last march point (ParaxialPrimaryEvolutionOperator.cpp:428, 430).  In a
    uniform tilted medium at normal incidence, with a Gaussian beam (a uniform
    beam has no transverse derivatives for the seam coefficients to act on):
    turning the duplicate plane's
    director about z (eps_zz unchanged, every other component changed) leaves
    every marched bulk plane bit for bit as it was, while tilting it further
    (eps_zz changed) moves them.  Reading the periodic neighbour instead of the
    duplicate fails the second check; reading more than eps_zz there fails the
    first.  Plane 0 is the entrance field after the input interface, which the
    duplicate points pass with their own permittivity (BPMIteration.cpp:87-122):
    there only the duplicate plane itself changes."""
    Nx, Ny, Nz = 9, 8, 4
    s = _settings((Nx, Ny, Nz), no="1.5", ne="1.7", beam_profile="GaussianBeam(0.3)")
    base = np.zeros((Nz, Ny, Nx, 3)) + _tilted(0.5, 0.4)

    def dup(n):
        lc = base.copy()
        if axis == "x":
            lc[:, :, Nx - 1, :] = n
        else:
            lc[:, Ny - 1, :, :] = n
        return lc

    _, b0 = oracle.propagate_bulk(s, base)
    _, b_turned = oracle.propagate_bulk(s, dup(_tilted(0.5, 1.5)))
    _, b_tilted = oracle.propagate_bulk(s, dup(_tilted(0.8, 0.4)))
    march = (slice(None),) * 3 + (slice(1, None),)
    assert np.array_equal(b_turned[march], b0[march])
    rel = np.linalg.norm(b_tilted[march] - b0[march]) / np.linalg.norm(b0[march])
    assert rel > 1e-6, rel
    # plane 0: equal at the march points, the duplicate plane through its own interface
    for b in (b_turned, b_tilted):
        assert np.array_equal(b[:, :, :, 0, :, :Ny - 1, :Nx - 1], b0[:, :, :, 0, :, :Ny - 1, :Nx - 1])
    edge = (Ellipsis, Nx - 1) if axis == "x" else (Ellipsis, Ny - 1, slice(None))
    assert not np.array_equal(b_turned[:, :, :, 0][edge], b0[:, :, :, 0][edge])
