"""GPU parity of the FP32 march (complex64 fields, FP64 coefficients) against
the FP64 CPU oracle on the input variants and launch shapes that the FP64
parity cases cover: every autotuner candidate shape (the FP32 solvers have
their own 16-point chunks and four-point gathers), Q-tensor input, a masked
droplet with a Gaussian beam at oblique incidence, mask-formula interface
weights, the caller's fields through isotropic layers with a focus shift,
Neumann depths 0 and 4, and meshes with axes of at most 3 points (no line
solves along them) or a single slab.  Tolerance: 1e-4 relative L2 per output
(wavelength, wavevector, polarisation).
"""
import json

import numpy as np
import pytest

import oracle
from nemaktis_amd import bpm_backend, fields
from test_gpu_parity import SOLVER_SHAPES, TOL32, _base, _formula_weights, _gpu, _random_case, _rel

pytestmark = pytest.mark.gpu

_want = {}


def _oracle_once(key, s, lc, **kw):
    """the oracle's output of a case, computed once per module"""
    if key not in _want:
        _want[key] = oracle.propagate(s, lc, **kw)
    return _want[key]


@pytest.mark.parametrize("tuning", SOLVER_SHAPES)
def test_line_solver_shapes_fp32(tuning):
    s, lc = _random_case((70, 66, 5), seed=17)
    want = _oracle_once("shapes", s, lc)
    err = _rel(_gpu(s, lc, precision=32, wl_batch=0, tuning=tuning), want)
    assert err < TOL32, err


def test_qtensor_field_fp32():
    dims = (16, 14, 5)
    n = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=9, corr_length=0.3)
    Q = 0.6 * (np.einsum("...i,...j->...ij", n, n) - np.eye(3) / 3)
    lc = np.stack([Q[..., 0, 0], Q[..., 1, 1], Q[..., 2, 2], Q[..., 0, 1], Q[..., 0, 2], Q[..., 1, 2]], -1)
    s = _base(dims, no="1.5", ne="1.7", field_type="Q-tensor", wavevectors=[(0.05, 0.1)])
    err = _rel(_gpu(s, lc, precision=32), oracle.propagate(s, lc))
    assert err < TOL32, err


def test_droplet_mask_gaussian_beam_fp32_graph():
    """masked droplet, Gaussian beam, normal and oblique incidence; replayed
    as a graph three times on one context (every replay equal)"""
    dims, sp = (24, 22, 8), (0.25, 0.25, 0.25)
    lc, mask, _ = fields.droplet(*dims, sp, 2.2)
    s = _base(dims, sp, wavevectors=[(0.0, 0.0), (0.1, 0.05)], no="1.5", ne="1.7", nhost="1.51",
              beam_profile="GaussianBeam(1.5)", na_objective=0.5)
    want = oracle.propagate(s, lc, mask=mask)
    shape = fields.output_shape(s)
    print("[gpu case] droplet mask, Gaussian beam, FP32 graph context", flush=True)
    with bpm_backend.Context(json.dumps(s), lc, int(np.prod(shape)), mask_vals=mask, precision=32, wl_batch=0,
                             use_graph=True) as ctx:
        outs = []
        for _ in range(3):
            ctx.propagate()
            outs.append(ctx.download(np.empty(shape, np.complex128)))
    assert all(np.array_equal(o, outs[0]) for o in outs)
    err = _rel(outs[0], want)
    assert err < TOL32, err


def test_droplet_mask_formula_weights_fp32():
    dims, sp = (20, 20, 9), (0.25, 0.25, 0.25)
    R = 1.9
    lc, mask, formula = fields.droplet(*dims, sp, R)
    s = _base(dims, sp, no="1.5", ne="1.7", nhost="1.51", na_objective=0.5)
    w = _formula_weights(s, lambda x, y, z: R * R - x * x - y * y - z * z, mask)
    assert w is not None
    want = oracle.propagate(s, lc, mask=mask, weights=w)
    err = _rel(_gpu(s, lc, mask=mask, formula=formula, precision=32), want)
    assert err < TOL32, err
    # the weights matter at this tolerance: without them the oracle's output moves by more
    assert _rel(oracle.propagate(s, lc, mask=mask), want) > 10 * TOL32


def test_iso_layers_focus_and_caller_fields_fp32():
    s, lc = _random_case((16, 16, 5), seed=5)
    s["Postprocessor settings"]["Screen output"].update({
        "Isotropic layer thicknesses": [1.0, 0.5], "Isotropic layer refractive indices": [1.45, 1.3],
        "Focalisation z-shift": 0.7})
    s["Physics settings"]["Initial conditions"]["Beam profile"] = "None"
    rng = np.random.default_rng(1)
    E_in = rng.standard_normal(fields.output_shape(s)) + 1j * rng.standard_normal(fields.output_shape(s))
    err = _rel(_gpu(s, lc, E_in=E_in, precision=32, wl_batch=0), oracle.propagate(s, lc, E=E_in))
    assert err < TOL32, err


@pytest.mark.parametrize("nw", [0, 4])
def test_neumann_depth_fp32(nw):
    s, lc = _random_case((14, 12, 4))
    s["Algorithm settings"]["Beam propagation"]["N Woodbury steps"] = nw
    err = _rel(_gpu(s, lc, precision=32, wl_batch=0), oracle.propagate(s, lc))
    assert err < TOL32, err


@pytest.mark.parametrize("dims", [(3, 20, 4), (18, 1, 4), (12, 10, 2), (2, 2, 3)])
def test_degenerate_meshes_fp32(dims):
    s = _base(dims, wavevectors=[(0.08, -0.12)], no="1.5", ne="1.7", na_objective=0.6)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=11, corr_length=0.3)
    err = _rel(_gpu(s, lc, precision=32), oracle.propagate(s, lc))
    assert err < TOL32, err
