"""GPU parity of the bulk (per-plane) output: the VTI file the backend writes
when "Bulk output" is active, read back and compared with the oracle's planes.

Tolerances: relative L2 per (wavelength, wavevector, polarisation) over all
planes, FP64 1e-10, FP32 1e-4 (as the output fields).
"""
import json
import os

import numpy as np
import pytest

import oracle
from nemaktis_amd import bpm_backend, fields, vti

pytestmark = pytest.mark.gpu


def _case(dims=(20, 18, 6), seed=3, folder="", name="bulk"):
    s = fields.make_settings(dims, (0.1, 0.12, 0.1), [0.55, 0.62], wavevectors=[(0.0, 0.0), (0.15, -0.1)],
                             no="1.5", ne="1.7+0.01/lambda^2", ne_imag="0.01", nin="1.2", nout="1.1",
                             na_objective=0.7)
    s["Algorithm settings"]["General"]["Results folder name"] = folder
    s["Postprocessor settings"]["Bulk output"] = {"Activate": True, "Base name": name}
    lc = fields.random_smooth(dims[0], dims[1], dims[2], (0.1, 0.12, 0.1), seed=seed, corr_length=0.3)
    return s, lc


def _compare(path, bulk_want, tol):
    got = vti.bulk_fields(path)
    nwl, nq = bulk_want.shape[:2]
    assert sorted(got) == [(w, q) for w in range(nwl) for q in range(nq)]
    worst = 0.0
    for (w, q), g in got.items():
        for pol in range(2):
            want = np.moveaxis(bulk_want[w, q, pol], 1, -1)          # (Nz, Ny, Nx, 2)
            err = np.linalg.norm(g[pol] - want) / np.linalg.norm(want)
            worst = max(worst, err)
    assert worst < tol, worst
    return worst


def test_bulk_output_fp64_matches_oracle(tmp_path):
    folder = str(tmp_path / "results")
    s, lc = _case(folder=folder)
    print("[gpu case] bulk fp64", flush=True)
    E = np.zeros(fields.output_shape(s), np.complex128)
    bpm_backend.run_backend_without_mask(json.dumps(s), lc, E, strict=True)
    E_want, bulk_want = oracle.propagate_bulk(s, lc)
    path = os.path.join(folder, "bulk.vti")
    _compare(path, bulk_want, 1e-10)
    geo, fd, pd = vti.read_vti(path)
    assert geo["dims"] == (20, 18, 6)
    np.testing.assert_allclose(geo["spacing"], (0.1, 0.12, 0.1))
    np.testing.assert_allclose(geo["origin"], (-0.95, -1.02, -0.25))
    np.testing.assert_array_equal(fd["lambda"], [0.55, 0.62])
    np.testing.assert_array_equal(fd["qx"], [0.0, 0.15])
    np.testing.assert_array_equal(fd["qy"], [0.0, -0.1])
    assert len(pd) == 2 * 2 * 2 * 2 and np.all(pd["E_imag_inputY_1_1"][..., 2] == 0)
    err = np.linalg.norm(E - E_want) / np.linalg.norm(E_want)
    assert err < 1e-10


def test_bulk_output_fp32_batched_graph_matches_oracle(tmp_path):
    s, lc = _case(dims=(24, 22, 5), seed=5, folder=str(tmp_path), name="b32")
    print("[gpu case] bulk fp32 graph", flush=True)
    with bpm_backend.Context(json.dumps(s), lc, int(np.prod(fields.output_shape(s))), precision=32, wl_batch=0,
                             use_graph=True) as ctx:
        ctx.propagate()
    _, bulk_want = oracle.propagate_bulk(s, lc)
    _compare(str(tmp_path / "b32.vti"), bulk_want, 1e-4)


def test_bulk_output_caller_fields_and_context_rewrites(tmp_path):
    # "Beam profile": "None": plane 0 is the caller's field as given; a resident context writes the
    # file on every propagation (the output replaces the uploaded fields, so they are uploaded again)
    s, lc = _case(dims=(16, 14, 4), seed=9, folder=str(tmp_path), name="ctx")
    s["Physics settings"]["Initial conditions"]["Beam profile"] = "None"
    rng = np.random.default_rng(2)
    shape = fields.output_shape(s)
    E_in = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    _, bulk_want = oracle.propagate_bulk(s, lc, E=E_in)
    path = str(tmp_path / "ctx.vti")
    with bpm_backend.Context(json.dumps(s), lc, int(np.prod(shape))) as ctx:
        ctx.upload_fields(E_in)
        ctx.propagate()
        _compare(path, bulk_want, 1e-10)
        os.remove(path)
        ctx.upload_fields(E_in)
        ctx.propagate()
        _compare(path, bulk_want, 1e-10)
    g = vti.bulk_fields(path)
    np.testing.assert_array_equal(g[(0, 0)][1, 0, ..., 0], E_in[0, 0, 2])


def test_bulk_output_unwritable_folder_is_reported(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    s, lc = _case(dims=(12, 10, 3), folder=str(blocker / "sub"))
    E = np.zeros(fields.output_shape(s), np.complex128)
    with pytest.raises(bpm_backend.BpmBackendError, match="results folder|bulk output file"):
        bpm_backend.run_backend_without_mask(json.dumps(s), lc, E, strict=True)


def test_bulk_output_of_a_serving_job(tmp_path):
    """A serving job with the bulk output active (its planes carved from the
    arena with the job's other buffers) writes the same file as the one-shot
    call, bit for bit, beside a job without it."""
    from nemaktis_amd import serving
    s, lc = _case(folder=str(tmp_path / "srv"), name="job")
    plain, lc2 = _case(seed=4)
    plain["Postprocessor settings"]["Bulk output"]["Activate"] = False
    with serving.Server(max_active=2) as srv:
        print("[gpu case] serving job with bulk output", flush=True)
        j = srv.submit(json.dumps(s), lc, precision=32)
        k = srv.submit(json.dumps(plain), lc2, precision=32)
        assert j.wait()["state"] == "done" and k.wait()["state"] == "done"
    s1 = json.loads(json.dumps(s))
    s1["Algorithm settings"]["General"]["Results folder name"] = str(tmp_path / "one")
    E = bpm_backend.propagate(json.dumps(s1), lc, precision=32, wl_batch=0)
    assert np.array_equal(j.E, E)
    a, b = vti.bulk_fields(str(tmp_path / "srv" / "job.vti")), vti.bulk_fields(str(tmp_path / "one" / "job.vti"))
    assert sorted(a) == sorted(b) and all(np.array_equal(a[key], b[key]) for key in a)


@pytest.mark.parametrize("precision", [64, 32])
def test_duplicate_plane_enters_the_march_only_through_eps_zz(tmp_path, precision):
    """The HIP march reads the duplicate column of the mesh only for the seam
    rule's 1/eps_zz (ParaxialPrimaryEvolutionOperator.cpp:428): turning the
    duplicate column's director about z (eps_zz unchanged) leaves every marched
    plane bit for bit as it was, tilting it further moves them as the oracle
    says; plane 0 changes only in the duplicate column (its own input
    interface)."""
    import math
    Nx, Ny, Nz = 9, 8, 4

    def tilted(theta, phi):
        return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])

    s = fields.make_settings((Nx, Ny, Nz), (0.1, 0.1, 0.1), [0.6], no="1.5", ne="1.7",
                             beam_profile="GaussianBeam(0.3)")
    base = np.zeros((Nz, Ny, Nx, 3)) + tilted(0.5, 0.4)
    runs = {}
    for name, n in (("base", None), ("turned", tilted(0.5, 1.5)), ("tilted", tilted(0.8, 0.4))):
        lc = base.copy()
        if n is not None:
            lc[:, :, Nx - 1, :] = n
        folder = str(tmp_path / name)
        s["Algorithm settings"]["General"]["Results folder name"] = folder
        s["Postprocessor settings"]["Bulk output"] = {"Activate": True, "Base name": "bulk"}
        print("[gpu case] duplicate column %s, precision %d" % (name, precision), flush=True)
        E = np.zeros(fields.output_shape(s), np.complex128)
        bpm_backend.run_backend_without_mask(json.dumps(s), lc, E, strict=True, precision=precision)
        runs[name] = (vti.bulk_fields(os.path.join(folder, "bulk.vti"))[(0, 0)], lc)   # (pol, Nz, Ny, Nx, 2)
    b0, bt, bs = runs["base"][0], runs["turned"][0], runs["tilted"][0]
    assert np.array_equal(bt[:, 1:], b0[:, 1:])
    assert np.array_equal(bt[:, 0, :, :Nx - 1], b0[:, 0, :, :Nx - 1])
    assert not np.array_equal(bt[:, 0, :, Nx - 1], b0[:, 0, :, Nx - 1])
    _, want = oracle.propagate_bulk(s, runs["tilted"][1])
    want = np.moveaxis(want[0, 0], 2, -1)                                            # (pol, Nz, Ny, Nx, 2)
    tol = 1e-10 if precision == 64 else 1e-4
    assert np.linalg.norm(bs - want) / np.linalg.norm(want) < tol
    assert np.linalg.norm(bs[:, 1:] - b0[:, 1:]) / np.linalg.norm(b0[:, 1:]) > 1e-4
