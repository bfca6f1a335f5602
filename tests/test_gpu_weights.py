"""Interface weights of a mask formula computed on the device (the masked
runs' path: one thread per column and slab, the formula compiled to a postfix program)
against the host computation (UtilitiesLib/VectorField.cpp:64-113 restated in
settings.cpp), which the CPU tests pin against the oracle and closed forms.

Tolerances: formulas of exact operations (+ - * /, x^2, comparisons, min,
max, avg, abs, sqrt, ternary) give the host's weights bit for bit; formulas
with transcendental functions may take a different bisection branch where the
formula is within an ulp of zero (the device and host math libraries round
differently), which moves a weight by at most the final bracket: at least
99.9 % of the weights equal, the rest within 2e-3.
"""
import json

import numpy as np
import pytest

from nemaktis_amd import _lib, fields

pytestmark = pytest.mark.gpu


def _weights(fn, s, formula, mask):
    Nx, Ny, Nz = s["Physics settings"]["Initial conditions"]["Mesh dimensions"]
    w = np.empty(2 * Nx * Ny * (Nz - 1))
    err = _lib.err_buf()
    m = np.ascontiguousarray(mask, dtype=np.float64).reshape(-1)
    rc = getattr(_lib.load(), fn)(json.dumps(s).encode(), formula.encode(), _lib.dptr(m), m.size, _lib.dptr(w),
                                  w.size, err, len(err))
    assert rc == 0, err.value
    return w.reshape(Nz - 1, Ny, Nx, 2), err.value.decode()


def _grid(dims, sp):
    x, y, z = fields.mesh_coords(*dims, *sp)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


EXACT = [
    ((100, 90, 30), (0.1, 0.1, 0.1), "1.6^2-x^2-y^2-z^2", lambda X, Y, Z: 1.6 ** 2 - X * X - Y * Y - Z * Z),
    ((64, 70, 21), (0.05, 0.05, 0.1), "0.13-z+0*x", lambda X, Y, Z: 0.13 - Z),
    ((40, 33, 17), (0.1, 0.1, 0.1),
     "(z>0.05 ? 0.9 : 1.1)^2-x^2-y^2-z^2+0*min(x,y)+0*avg(x,y,z)+0*abs(sqrt(x^2+1))",
     lambda X, Y, Z: np.where(Z > 0.05, 0.9, 1.1) ** 2 - X * X - Y * Y - Z * Z),
]


@pytest.mark.parametrize("ci", range(len(EXACT)))
def test_device_weights_equal_host_weights(ci):
    dims, sp, formula, fn = EXACT[ci]
    print("[gpu case] device weights", dims, formula, flush=True)
    s = fields.make_settings(dims, sp, [0.6])
    mask = fn(*_grid(dims, sp))
    wd, nd = _weights("nmk_bpm_interface_weights_device", s, formula, mask)
    wh, nh = _weights("nmk_bpm_interface_weights", s, formula, mask)
    assert nd == nh == ""
    assert np.array_equal(wd, wh)
    assert np.sum(wd[..., 0] != 0.5) > 100


def test_device_weights_with_transcendental_functions():
    dims, sp = (96, 80, 25), (0.1, 0.1, 0.1)
    formula = "z-0.3*sin(2*x)*cos(1.5*y)+0.1*exp(-x^2)-0.05*tanh(y)"
    X, Y, Z = _grid(dims, sp)
    mask = Z - 0.3 * np.sin(2 * X) * np.cos(1.5 * Y) + 0.1 * np.exp(-X * X) - 0.05 * np.tanh(Y)
    s = fields.make_settings(dims, sp, [0.6])
    print("[gpu case] device weights, transcendental", flush=True)
    wd, nd = _weights("nmk_bpm_interface_weights_device", s, formula, mask)
    wh, nh = _weights("nmk_bpm_interface_weights", s, formula, mask)
    assert nd == nh == ""
    same = wd == wh
    assert same.mean() >= 0.999, same.mean()
    assert np.abs(wd - wh).max() <= 2e-3
    assert np.sum(wd[..., 0] != 0.5) >= dims[0] * dims[1] // 2


def test_device_weights_mismatch_gives_the_notice():
    dims, sp = (30, 20, 8), (0.2, 0.2, 0.2)
    s = fields.make_settings(dims, sp, [0.6])
    wd, nd = _weights("nmk_bpm_interface_weights_device", s, "0.05-z", np.ones(dims[::-1]))
    assert np.all(wd == 0.5)
    assert nd == "Mask values do not corresponds to the given formula, will ignore interpolation correction"
