"""The bpm-solver command line on the CPU: settings files, the VTI reader in
every encoding VTK writes for image data, the LC field built from a VTI file
(basis conventions, Q-tensor, domain mask), usage and errors."""
import base64
import ctypes
import json
import zlib

import numpy as np
import pytest

from nemaktis_amd import _lib, cli, vti


# ---- VTI files in VTK's encodings (layout of the VTK XML file-format description) --------------------------

def _blocks(data, htype, compress, block=64):
    """(header bytes, payload bytes) of one array: plain [nbytes] + data, or the compressed block table
    [nblocks, blocksize, last partial size, compressed sizes...] + zlib blocks."""
    h = np.dtype(htype)
    if not compress:
        return np.array([len(data)], h).tobytes(), data
    parts = [data[i:i + block] for i in range(0, len(data), block)] or [b""]
    comp = [zlib.compress(p) for p in parts]
    last = len(parts[-1]) if len(parts[-1]) != block else 0
    head = np.array([len(parts), block, last] + [len(c) for c in comp], h).tobytes()
    return head, b"".join(comp)


def _write(path, arrays, dims, mode, htype="<u4", compress=False):
    """arrays: {name: (values float64 (npts, nc) or (npts,))}; mode: ascii | binary | raw | base64."""
    Nx, Ny, Nz = dims
    hdr = "UInt64" if np.dtype(htype).itemsize == 8 else "UInt32"
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    out = ['<?xml version="1.0"?>', '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
           'header_type="%s"%s>' % (hdr, comp),
           '<ImageData WholeExtent="0 %d 0 %d 0 %d" Origin="0 0 0" Spacing="0.1 0.2 0.3">' % (Nx - 1, Ny - 1, Nz - 1),
           '<Piece Extent="0 %d 0 %d 0 %d">' % (Nx - 1, Ny - 1, Nz - 1), "<PointData>"]
    appended, off = [], 0
    for name, v in arrays.items():
        v = np.asarray(v, "<f8")
        nc = 1 if v.ndim == 1 else v.shape[1]
        attrs = 'type="Float64" Name="%s" NumberOfComponents="%d"' % (name, nc)
        head, payload = _blocks(v.tobytes(), htype, compress)
        if mode == "ascii":
            out.append('<DataArray %s format="ascii">%s</DataArray>' % (attrs, " ".join(repr(float(x)) for x in v.ravel())))
        elif mode == "binary":
            text = (base64.b64encode(head) + base64.b64encode(payload)).decode() if compress else \
                base64.b64encode(head + payload).decode()
            out.append('<DataArray %s format="binary">\n  %s\n</DataArray>' % (attrs, text))
        else:
            if mode == "raw":
                blob = head + payload
            else:
                blob = (base64.b64encode(head) + base64.b64encode(payload)) if compress else \
                    base64.b64encode(head + payload)
            out.append('<DataArray %s format="appended" offset="%d"/>' % (attrs, off))
            appended.append(blob)
            off += len(blob)
    out += ["</PointData>", "</Piece>", "</ImageData>"]
    body = "\n".join(out).encode()
    if appended:
        body += ('\n<AppendedData encoding="%s">\n_' % ("raw" if mode == "raw" else "base64")).encode()
        body += b"".join(appended) + b"\n</AppendedData>"
    body += b"\n</VTKFile>\n"
    with open(path, "wb") as f:
        f.write(body)


@pytest.mark.parametrize("mode,htype,compress", [
    ("ascii", "<u4", False), ("binary", "<u4", False), ("binary", "<u8", True), ("raw", "<u4", False),
    ("raw", "<u8", True), ("base64", "<u8", False), ("base64", "<u4", True)])
def test_vti_reader_encodings(tmp_path, mode, htype, compress):
    dims = (5, 4, 3)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((60, 3))
    b = rng.standard_normal(60)
    p = tmp_path / "f.vti"
    _write(p, {"n": a, "S": b}, dims, mode, htype, compress)
    geo, _, pd = vti.read_vti(p)
    assert geo["dims"] == dims and geo["spacing"] == (0.1, 0.2, 0.3)
    np.testing.assert_array_equal(pd["n"], a.reshape(3, 4, 5, 3))
    np.testing.assert_array_equal(pd["S"], b.reshape(3, 4, 5))


def test_vti_writer_roundtrip(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((24, 3))
    p = tmp_path / "w.vti"
    vti.write_vti(p, (4, 3, 2), (0.1, 0.1, 0.2), (-0.15, -0.1, -0.1), {"lambda": [0.5, 0.6]}, {"E": a})
    geo, fd, pd = vti.read_vti(p)
    assert geo["dims"] == (4, 3, 2) and geo["origin"] == (-0.15, -0.1, -0.1)
    np.testing.assert_array_equal(fd["lambda"], [0.5, 0.6])
    np.testing.assert_array_equal(pd["E"], a.reshape(2, 3, 4, 3))


# ---- settings files ------------------------------------------------------------------------------------------

def test_minify_drops_blank_and_comment_lines(tmp_path):
    p = tmp_path / "s.json"
    p.write_text('{\n  # comment\n\n   \t# indented comment\n  "a": 1 # not a comment line\n}\n')
    assert cli.minify_json(str(p)) == '{\n  "a": 1 # not a comment line\n}\n'


def test_default_settings_file_parses_with_the_backend_schema(tmp_path):
    p = tmp_path / "default.json"
    assert cli.main(["-c", str(p)]) == 0
    s = json.loads(cli.minify_json(str(p)))
    ic = s["Physics settings"]["Initial conditions"]
    assert ic["Basis convention"] == "XYZ" and s["Algorithm settings"]["General"]["LC field type"] == "Director"
    ic["Mesh dimensions"], ic["Mesh spacings"] = [8, 8, 4], [0.1, 0.1, 0.1]
    lib = _lib.load()
    err = _lib.err_buf()
    v = ctypes.c_int64()
    opts = _lib.options()
    rc = lib.nmk_bpm_estimate_footprint(json.dumps(s).encode(), 8 * 8 * 4 * 3, 1, ctypes.byref(opts),
                                        ctypes.byref(v), err, len(err))
    assert rc == 0, err.value
    wl, qs = cli._wavelengths_wavevectors(s["Physics settings"]["Coefficients"])
    assert wl == [0.6] and qs == [(0.0, 0.0)]


def test_wavelength_spectrum_and_condenser_rings():
    wl, qs = cli._wavelengths_wavevectors({"Mean wavelength": 0.6, "Spectral FWHM": 0.2, "N wavelengths": 3,
                                           "Condenser numerical aperture": 0.2,
                                           "N radial illumination directions": 2})
    np.testing.assert_allclose(wl, [0.5, 0.6, 0.7])
    assert len(qs) == 7
    # the reference steps the ring angle by pi/3 whatever the ring (PhysicsSettings.cpp:74-76)
    np.testing.assert_allclose(qs[2], (0.2 * np.cos(np.pi / 3), 0.2 * np.sin(np.pi / 3)))


# ---- LC field from a VTI file ----------------------------------------------------------------------------------

def _director_file(tmp_path, dims, n, extra=None):
    p = tmp_path / "lc.vti"
    arrays = {"n": n.reshape(-1, 3)}
    if extra:
        arrays.update(extra)
    _write(p, arrays, dims, "raw", "<u8", True)
    return str(p)


@pytest.mark.parametrize("basis", ["XYZ", "YZX", "ZXY"])
def test_director_basis_conventions(tmp_path, basis):
    dims = (4, 3, 2)                          # file x, y, z sizes
    rng = np.random.default_rng(3)
    n = rng.standard_normal((2, 3, 4, 3))     # file array (z, y, x, component)
    lc, mask, mdims, msp = cli.lc_field_from_vti(_director_file(tmp_path, dims, n), "Director", basis)
    I = {"XYZ": (0, 1, 2), "YZX": (1, 2, 0), "ZXY": (2, 0, 1)}[basis]
    assert mdims == [dims[i] for i in I]
    np.testing.assert_allclose(msp, [(0.1, 0.2, 0.3)[i] for i in I])
    assert lc.shape == (mdims[2], mdims[1], mdims[0], 3)
    # VTIReader.cpp:104-124: i_out = idx[I1] + dim[I1] (idx[I2] + dim[I2] idx[I3]), components (I1, I2, I3)
    for ix in range(4):
        for iy in range(3):
            for iz in range(2):
                idx = (ix, iy, iz)
                np.testing.assert_array_equal(lc[idx[I[2]], idx[I[1]], idx[I[0]]], n[iz, iy, ix][list(I)])
    assert np.all(mask == 1)


def test_q_tensor_from_director_and_order(tmp_path):
    dims = (3, 2, 2)
    rng = np.random.default_rng(4)
    n = rng.standard_normal((2, 2, 3, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    S = rng.uniform(0, 0.5, (2, 2, 3))
    lc, _, _, _ = cli.lc_field_from_vti(_director_file(tmp_path, dims, n, {"S": S.reshape(-1)}), "Q-tensor", "XYZ")
    f = 1 - S
    # Q = (1 - S) (3 n n^T - I) / 2 with the off-diagonal entries xy, xz, yz
    Q = 0.5 * f[..., None, None] * (3 * n[..., :, None] * n[..., None, :] - np.eye(3))
    np.testing.assert_allclose(lc, np.stack([Q[..., 0, 0], Q[..., 1, 1], Q[..., 2, 2], Q[..., 0, 1],
                                             Q[..., 0, 2], Q[..., 1, 2]], -1), atol=1e-15)
    q = rng.standard_normal((2, 2, 3, 6))
    p = tmp_path / "q.vti"
    _write(p, {"Q": q.reshape(-1, 6)}, dims, "binary", "<u4", False)
    lc, _, _, _ = cli.lc_field_from_vti(str(p), "Q-tensor", "XYZ")
    np.testing.assert_array_equal(lc, q)


def test_domain_mask_and_errors(tmp_path):
    dims = (3, 3, 2)
    n = np.zeros((2, 3, 3, 3))
    n[..., 2] = 1
    n[0, 1, 1] = 0
    lc, mask, _, _ = cli.lc_field_from_vti(_director_file(tmp_path, dims, n), "Director", "XYZ")
    assert mask[0, 1, 1] == -1 and (mask == 1).sum() == 17
    with pytest.raises(cli.CliError, match="must be a path to a vti file"):
        cli.lc_field_from_vti("lc.vtk", "Director", "XYZ")
    with pytest.raises(cli.CliError, match="does not exists"):
        cli.lc_field_from_vti(str(tmp_path / "none.vti"), "Director", "XYZ")
    n[1, 2, 2, 0] = np.nan
    with pytest.raises(cli.CliError, match="NaN value"):
        cli.lc_field_from_vti(_director_file(tmp_path, dims, n), "Director", "XYZ")
    p = tmp_path / "nothing.vti"
    _write(p, {"m": np.zeros(18)}, dims, "raw")
    with pytest.raises(cli.CliError, match='array "n"'):
        cli.lc_field_from_vti(str(p), "Director", "XYZ")


def test_usage_and_errors(capsys, tmp_path):
    assert cli.main(["-h"]) == 0
    assert "-x FILE" in capsys.readouterr().out
    assert cli.main([]) == 0 and "Syntax error" in capsys.readouterr().out
    assert cli.main(["-x", str(tmp_path / "missing.json")]) == 1
    assert "Could not open the specified setting file" in capsys.readouterr().err
