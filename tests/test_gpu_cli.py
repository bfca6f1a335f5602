"""GPU end-to-end runs of the bpm-solver command line: a commented settings
file and a director VTI file in, the screen and bulk VTI files out, compared
with the oracle on the same LC field and domain mask.

The director is built on the solver's axes and written to the file on the
axes of the basis convention, so the file-to-solver mapping is checked end to
end.  Tolerances: relative L2 per output set, FP64 1e-10, FP32 1e-4.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from nemaktis_amd import cli, vti

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_BASIS = {"XYZ": (0, 1, 2), "YZX": (1, 2, 0), "ZXY": (2, 0, 1)}


def _director(Nx, Ny, Nz, seed):
    """Smooth random director on the solver's axes (Nz, Ny, Nx, 3) with a host-medium hole (zero director)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    th = 0.8 * np.sin(0.4 * x + rng.uniform(0, 6)) * np.cos(0.3 * y + rng.uniform(0, 6)) + 0.2 * z
    ph = 1.1 * np.cos(0.35 * x - 0.25 * y + rng.uniform(0, 6))
    n = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    hole = (x - Nx / 2) ** 2 + (y - Ny / 2) ** 2 < (min(Nx, Ny) / 4) ** 2
    n[hole & (z >= 1) & (z <= Nz - 2)] = 0
    return n


def _write_director_file(path, n, spacings, basis):
    """Writes the solver-axes director n (Nz, Ny, Nx, 3) as the file that the basis convention maps onto it."""
    I = _BASIS[basis]
    # solver axis a is file axis I[a]; file array axes (z, y, x) are file axes (2, 1, 0)
    solver_of_file = [I.index(f) for f in range(3)]               # file axis f -> solver axis
    arr = np.transpose(n, [2 - solver_of_file[2 - k] for k in range(3)] + [3])
    comps = np.empty_like(arr)
    for a in range(3):
        comps[..., I[a]] = arr[..., a]
    fdims = [0, 0, 0]
    fsp = [0.0, 0.0, 0.0]
    for a, d in enumerate((n.shape[2], n.shape[1], n.shape[0])):
        fdims[I[a]] = d
        fsp[I[a]] = spacings[a]
    assert comps.shape[:3] == (fdims[2], fdims[1], fdims[0])
    vti.write_vti(path, tuple(fdims), tuple(fsp), (0.0, 0.0, 0.0), {}, {"n": comps.reshape(-1, 3)})


def _settings_file(path, lc_file, folder, basis, nwl, nr):
    text = cli.DEFAULT_SETTINGS
    for old, new in (('"LC field file":\t""', '"LC field file":\t"%s"' % lc_file),
                     ('"Results folder name":\t"results/"', '"Results folder name":\t"%s"' % folder),
                     ('"Basis convention":\t"XYZ"', '"Basis convention":\t"%s"' % basis),
                     ('"nhost": \t\t\t\t"1.5"', '"nhost": \t\t\t\t"1.33"'),
                     ('"N wavelengths":\t1', '"N wavelengths":\t%d' % nwl),
                     ('"Condenser numerical aperture": 0', '"Condenser numerical aperture": 0.2'),
                     ('"N radial illumination directions": 1', '"N radial illumination directions": %d' % nr),
                     ('"Activate":\t\t\t\t\t\t\t\tfalse', '"Activate":\t\t\t\t\t\t\t\ttrue')):
        assert old in text
        text = text.replace(old, new)
    with open(path, "w") as f:
        f.write(text)


def _screen_fields(path, nwl, nq):
    geo, fd, pd = vti.read_vti(path)
    Nx, Ny, _ = geo["dims"]
    E = np.zeros((nwl, nq, 4, Ny, Nx), np.complex128)
    for w in range(nwl):
        for q in range(nq):
            for pol, tag in enumerate("XY"):
                c = pd["E_real_input%s_%d_%d" % (tag, w, q)][0] + 1j * pd["E_imag_input%s_%d_%d" % (tag, w, q)][0]
                E[w, q, 2 * pol], E[w, q, 2 * pol + 1] = c[..., 0], c[..., 1]
    return geo, fd, E


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _run_case(tmp_path, basis, argv_extra, in_process, nwl=2, nr=2, dims=(18, 16, 6), sp=(0.1, 0.12, 0.1)):
    n = _director(dims[0], dims[1], dims[2], seed=len(basis) + len(argv_extra))
    lc_file = str(tmp_path / "director.vti")
    _write_director_file(lc_file, n, sp, basis)
    folder = str(tmp_path / "out")
    settings = str(tmp_path / "settings.json")
    _settings_file(settings, lc_file, folder, basis, nwl, nr)
    print("[gpu case] cli basis=%s %s" % (basis, " ".join(argv_extra)), flush=True)
    if in_process:
        assert cli.main(["-x", settings] + argv_extra) == 0
    else:
        r = subprocess.run([sys.executable, "-m", "nemaktis_amd.cli", "-x", settings] + argv_extra, cwd=ROOT,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
    # the oracle's input: the solver-axes field, the domain where it is non-zero, the mesh filled in
    import json
    s = json.loads(cli.minify_json(settings))
    ic = s["Physics settings"]["Initial conditions"]
    ic["Mesh dimensions"], ic["Mesh spacings"] = list(dims), list(sp)
    mask = np.where(np.linalg.norm(n, axis=-1) > 1e-8, 1.0, -1.0)
    assert (mask < 0).any() and (mask > 0).any()
    return s, n, mask, folder


def test_cli_fp64_screen_and_bulk_match_oracle(tmp_path):
    s, n, mask, folder = _run_case(tmp_path, "YZX", [], in_process=True)
    nq = 7
    geo, fd, E = _screen_fields(os.path.join(folder, "screen.vti"), 2, nq)
    assert geo["dims"] == (18, 16, 1)
    np.testing.assert_allclose(geo["spacing"], (0.1, 0.12, 0.0))
    np.testing.assert_allclose(fd["lambda"], [0.5, 0.7])
    assert len(fd["qx"]) == nq
    E_want, bulk_want = oracle.propagate_bulk(s, n, mask=mask)
    assert _rel(E, E_want) < 1e-10
    got = vti.bulk_fields(os.path.join(folder, "bulk.vti"))
    assert sorted(got) == [(w, q) for w in range(2) for q in range(nq)]
    worst = max(_rel(g[pol], np.moveaxis(bulk_want[w, q, pol], 1, -1)) for (w, q), g in got.items()
                for pol in range(2))
    assert worst < 1e-10, worst


def test_cli_fp32_batched_graph_subprocess_matches_oracle(tmp_path):
    s, n, mask, folder = _run_case(tmp_path, "ZXY", ["--precision", "32", "--batch", "--graph"],
                                   in_process=False, nwl=3, nr=1, dims=(20, 14, 5), sp=(0.11, 0.1, 0.12))
    _, _, E = _screen_fields(os.path.join(folder, "screen.vti"), 3, 1)
    E_want = oracle.propagate(s, n, mask=mask)
    assert _rel(E, E_want) < 1e-4
