"""The no-allocation entry point nmk_bpm_run_noalloc against the existing
entry point nmk_bpm_run on the same inputs, through the C ABI: two calls in a
row on one GPU (the second with a smaller input, so the scratch array holds
the first call's contents), FP64 and FP32, with and without a mask (its
interface weights precomputed by nmk_bpm_interface_weights), plus the
refusal of a run larger than the scratch.  Both entry points run the same
kernels with the same launch shapes, so the outputs are equal bit for bit
(the FP64 one is also checked against the oracle, 1e-10).  The test's device
arrays come from the HIP runtime the library itself links (libamdhip64.so.7),
called through ctypes: test plumbing only."""
import ctypes
import json

import numpy as np
import pytest

import oracle
from nemaktis_amd import _lib, bpm_backend, fields

pytestmark = pytest.mark.gpu


_hip = None


def _runtime():
    global _hip
    if _hip is None:
        _lib.load()  # loads libamdhip64.so.7 with the library; the same instance is returned below
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


class _DevArray:
    """device copy of a float64 numpy array (hipMalloc + hipMemcpy)"""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a, np.float64).reshape(-1)
        self.n = self.a.size
        self.p = ctypes.c_void_p()
        assert _runtime().hipMalloc(ctypes.byref(self.p), self.a.nbytes) == 0
        assert _runtime().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert _runtime().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            _runtime().hipFree(self.p)
            self.p = ctypes.c_void_p()


def _noalloc(s, lc, mask=None, weights=None, precision=64, E_in=None):
    lib = _lib.load()
    shape = fields.output_shape(s)
    E0 = np.zeros(shape, np.complex128) if E_in is None else np.asarray(E_in, np.complex128)
    arrs = [_DevArray(lc), None if mask is None else _DevArray(mask), None if weights is None else _DevArray(weights),
            _DevArray(E0.reshape(-1).view(np.float64))]
    lc_d, m_d, w_d, E_d = arrs
    try:
        err = _lib.err_buf()
        rc = lib.nmk_bpm_run_noalloc(json.dumps(s).encode(), lc_d.p, lc_d.n, None if m_d is None else m_d.p,
                                     0 if m_d is None else m_d.n, None if w_d is None else w_d.p,
                                     0 if w_d is None else w_d.n, E_d.p, int(np.prod(shape)), precision, None,
                                     err, len(err))
        if rc:
            raise bpm_backend.BpmBackendError("rc %d: %s" % (rc, err.value.decode()))
        return E_d.get().view(np.complex128).reshape(shape)
    finally:
        for a in arrs:
            if a is not None:
                a.free()


def _weights(s, formula, mask):
    Nx, Ny, Nz = s["Physics settings"]["Initial conditions"]["Mesh dimensions"]
    w = np.empty(2 * Nx * Ny * (Nz - 1))
    err = _lib.err_buf()
    m = np.ascontiguousarray(mask, np.float64).reshape(-1)
    assert _lib.load().nmk_bpm_interface_weights(json.dumps(s).encode(), formula.encode(), _lib.dptr(m), m.size,
                                                 _lib.dptr(w), w.size, err, len(err)) == 0
    return w


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_equals_existing_entry_point_two_calls():
    # first call: walk-off, oblique incidence, 2 wavelengths x 2 wavevectors, FP64
    dims = (40, 36, 6)
    s = fields.make_settings(dims, (0.1, 0.1, 0.1), [0.55, 0.62], wavevectors=[(0.0, 0.0), (0.12, -0.08)],
                             no="1.5", ne="1.7", ne_imag="0.01", na_objective=0.6)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=3, corr_length=0.3)
    print("[gpu case] noalloc call 1: 40x36x6 FP64", flush=True)
    got = _noalloc(s, lc)
    want = bpm_backend.propagate(json.dumps(s), lc)
    assert np.array_equal(got, want)
    assert _rel(got, oracle.propagate(s, lc)) < 1e-10
    # second call, smaller: masked droplet with formula weights, FP32
    dims2, sp2 = (24, 20, 7), (0.25, 0.25, 0.25)
    lc2, mask2, formula2 = fields.droplet(*dims2, sp2, 1.6)
    s2 = fields.make_settings(dims2, sp2, [0.6], no="1.5", ne="1.7", nhost="1.51", na_objective=0.5)
    w2 = _weights(s2, formula2, mask2)
    print("[gpu case] noalloc call 2: 24x20x7 FP32 masked", flush=True)
    got2 = _noalloc(s2, lc2, mask=mask2, weights=w2, precision=32)
    want2 = bpm_backend.propagate(json.dumps(s2), lc2, mask_vals=mask2, mask_formula=formula2, precision=32)
    assert np.array_equal(got2, want2)
    # a third call repeating the first: the scratch's leftovers change nothing
    print("[gpu case] noalloc call 3: 40x36x6 FP64 again", flush=True)
    assert np.array_equal(_noalloc(s, lc), want)


def test_noalloc_caller_fields_fp64():
    dims = (18, 16, 5)
    s = fields.make_settings(dims, (0.1, 0.1, 0.1), [0.6], wavevectors=[(0.05, 0.1)], no="1.5", ne="1.7",
                             beam_profile="None", na_objective=0.6)
    lc = fields.random_smooth(*dims, (0.1, 0.1, 0.1), seed=5, corr_length=0.3)
    rng = np.random.default_rng(4)
    shape = fields.output_shape(s)
    E_in = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    print("[gpu case] noalloc caller fields", flush=True)
    got = _noalloc(s, lc, E_in=E_in)
    want = bpm_backend.propagate(json.dumps(s), lc, E_in=E_in)
    assert np.array_equal(got, want)


def test_noalloc_refuses_runs_larger_than_its_scratch():
    dims = (1024, 1024, 3)
    s = fields.make_settings(dims, (0.1, 0.1, 0.1), [0.6])
    lc = np.zeros(dims[::-1] + (3,))
    lc[..., 2] = 1
    print("[gpu case] noalloc oversize refusal", flush=True)
    with pytest.raises(bpm_backend.BpmBackendError, match="rc 3: the run needs .* bytes of device scratch"):
        _noalloc(s, lc)
