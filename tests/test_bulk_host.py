"""Bulk (per-plane) output on the CPU: pins of the oracle's bulk planes by
closed forms, the footprint of a bulk run, and the VTI reader."""
import ctypes
import json

import numpy as np
import pytest

import oracle
from nemaktis_amd import _lib, fields, vti


def _iso(dims, q=(0.0, 0.0), n="1.5", nin="1.0"):
    # director along z: the transverse permittivity is no^2 I, an isotropic medium for the march
    s = fields.make_settings(dims, (0.1, 0.1, 0.1), [0.6], wavevectors=[q], no=n, ne="1.7", nhost=n, nin=nin,
                             nout=n, na_objective=1.0)
    lc = np.zeros((dims[2], dims[1], dims[0], 3))
    lc[..., 2] = 1.0
    return s, lc


@pytest.mark.parametrize("q", [(0.0, 0.0), (0.1, -0.05)])
def test_bulk_planes_of_a_homogeneous_isotropic_slab(q):
    # uniform beam in a transversely isotropic medium: the phase operator and the ADI step are the identity, the
    # shifts move a uniform field onto itself, so plane k is plane 0 times exp(2 i k theta),
    # theta = -k0 dz (qx^2 + qy^2) / (4 n); plane 0 is the textbook transmission 2 nin / (nin + n)
    n, nin = 1.5, 1.0
    s, lc = _iso((8, 6, 5), q)
    _, bulk = oracle.propagate_bulk(s, lc)
    k0dz = 2 * np.pi / 0.6 * 0.1
    theta = -k0dz * (q[0] ** 2 + q[1] ** 2) / (4 * n)
    t = 2 * nin / (nin + n)
    for pol in range(2):
        want0 = np.zeros((2, 6, 8), complex)
        want0[pol] = t
        for k in range(5):
            np.testing.assert_allclose(bulk[0, 0, pol, k], want0 * np.exp(2j * k * theta), atol=1e-13)


def test_last_bulk_plane_through_the_output_interface_is_the_output():
    # the output fields are the last plane after the output interface (BPMIteration.cpp:170-176)
    s, lc = fields.make_settings((14, 12, 5), (0.1, 0.1, 0.1), [0.55], wavevectors=[(0.1, 0.05)], no="1.5",
                                 ne="1.7", nout="1.2", na_objective=1.0), fields.random_smooth(14, 12, 5,
                                                                                               (0.1, 0.1, 0.1),
                                                                                               seed=3,
                                                                                               corr_length=0.3)
    E = oracle.propagate(s, lc, apply_filter=False)
    _, bulk = oracle.propagate_bulk(s, lc)
    sess = oracle.Session(s, lc)
    try:
        sess.set_wavelength(0)
        e = sess.eps_plane(4).reshape(12, 14, 12)
    finally:
        sess.close()
    nxx, nyy, nxy = e[..., 9], e[..., 10], e[..., 11]
    n = 1.2
    det = (n + nxx) * (n + nyy) - nxy * nxy
    cxx, cyy, cxy = 1 - n * (n + nyy) / det, 1 - n * (n + nxx) / det, n * nxy / det
    for pol in range(2):
        a, b = bulk[0, 0, pol, 4, 0], bulk[0, 0, pol, 4, 1]
        np.testing.assert_allclose(E[0, 0, 2 * pol], 2 * (cxx * a + cxy * b), atol=1e-12)
        np.testing.assert_allclose(E[0, 0, 2 * pol + 1], 2 * (cyy * b + cxy * a), atol=1e-12)


def test_footprint_counts_the_bulk_planes():
    s, _ = _iso((20, 16, 7))
    lib = _lib.load()
    err = _lib.err_buf()

    def fp(activate):
        s["Postprocessor settings"]["Bulk output"]["Activate"] = activate
        v = ctypes.c_int64()
        opts = _lib.options()
        assert lib.nmk_bpm_estimate_footprint(json.dumps(s).encode(), 20 * 16 * 7 * 3, 0, ctypes.byref(opts),
                                              ctypes.byref(v), err, len(err)) == 0
        return v.value

    extra = fp(True) - fp(False)
    planes = 2 * 7 * 2 * 20 * 16 * 16      # pols x planes x comps x points x complex128
    shifted = 2 * 2 * 19 * 15 * 16         # one march-grid state of both pols (FP64)
    assert planes + shifted <= extra <= planes + shifted + 2 * 256


def test_vti_reader_roundtrip(tmp_path):
    # a hand-written raw-appended file in the layout the writer produces
    Nx, Ny, Nz = 3, 2, 2
    data = np.arange(Nx * Ny * Nz * 3, dtype=np.float64)
    lam = np.array([0.5, 0.6])
    head = ('<?xml version="1.0"?>\n<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64">\n<ImageData WholeExtent="0 2 0 1 0 1" Origin="-0.1 -0.05 -0.05" '
            'Spacing="0.1 0.1 0.1" Direction="1 0 0 0 1 0 0 0 1">\n<FieldData>\n'
            '<DataArray type="Float64" Name="lambda" NumberOfTuples="2" format="appended" offset="0"/>\n'
            '</FieldData>\n<Piece Extent="0 2 0 1 0 1">\n<PointData>\n'
            '<DataArray type="Float64" Name="E_real_inputX_0_0" NumberOfComponents="3" format="appended" '
            'offset="24"/>\n</PointData>\n</Piece>\n</ImageData>\n<AppendedData encoding="raw">\n_')
    p = tmp_path / "t.vti"
    with open(p, "wb") as f:
        f.write(head.encode())
        f.write(np.uint64(16).tobytes() + lam.tobytes())
        f.write(np.uint64(data.nbytes).tobytes() + data.tobytes())
        f.write(b"\n</AppendedData>\n</VTKFile>\n")
    geo, fd, pd = vti.read_vti(p)
    assert geo["dims"] == (3, 2, 2) and geo["spacing"] == (0.1, 0.1, 0.1)
    np.testing.assert_array_equal(fd["lambda"], lam)
    np.testing.assert_array_equal(pd["E_real_inputX_0_0"], data.reshape(Nz, Ny, Nx, 3))
