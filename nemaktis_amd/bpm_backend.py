"""Drop-in replacement of the Nemaktis ``bpm_backend`` module on MI355X.

The two functions keep the reference's names, argument order and array
semantics (BeamPropagationSolver/src/PythonBackend/bpm_backend.h:4-13,
SWIG typemaps bpm_backend.i:16-18, caller
HighLevelPythonInterface/nemaktis/light_propagator.py:225-228):

* ``lc_field_vals`` / ``mask_vals``: any array convertible to contiguous
  float64 (SWIG ``IN_ARRAY1`` semantics: converted and copied);
* ``E_field_vals``: must already be a C-contiguous complex128 array
  (``INPLACE_ARRAY_FLAT``); it is overwritten in place.

Error behaviour follows the reference by default: the message is printed on
stderr in the reference's frame and the call returns normally
(PythonBackend/bpm_backend.cpp:48-82).  Set ``NEMAKTIS_BPM_STRICT=1`` (or
pass ``strict=True``) to raise ``BpmBackendError`` instead.  Array-type
violations raise ``TypeError`` as the SWIG typemaps do.

Precision and batching options (``precision=32``, ``wl_batch=0``) are
extensions; the defaults run the FP64 path one wavelength at a time, as the
reference does.  One-shot calls launch their kernels eagerly: a graph captured
for one march would be replayed once.  Graph replay belongs to ``Context``
(``use_graph=True``), which keeps its buffers and graphs between
propagations.
"""
import ctypes
import os
import sys

import numpy as np

from . import _lib

__all__ = ["run_backend_without_mask", "run_backend_with_mask", "BpmBackendError", "propagate", "Context"]


class BpmBackendError(RuntimeError):
    pass


def _strict(strict):
    if strict is not None:
        return bool(strict)
    return os.environ.get("NEMAKTIS_BPM_STRICT", "0") not in ("", "0", "false", "False")


def _check_E(E):
    if not isinstance(E, np.ndarray):
        raise TypeError("array of type 'complex128' required")
    if E.dtype != np.complex128:
        raise TypeError("Array of type 'complex128' required.  Array of type '%s' given" % E.dtype)
    if not E.flags["C_CONTIGUOUS"]:
        raise TypeError("Array must be contiguous.  A non-contiguous array was given")
    if not E.flags["WRITEABLE"]:
        raise TypeError("Array must be writeable")


def _report(rc, err, strict):
    if rc == 0:
        # a notice the reference prints to stdout on success (mask formula
        # that disagrees with the mask, UtilitiesLib/VectorField.cpp:107-111)
        if err.value:
            sys.stdout.write(err.value.decode("utf-8", "replace") + "\n")
            sys.stdout.flush()
        return
    msg = err.value.decode("utf-8", "replace")
    if _strict(strict):
        raise BpmBackendError(msg)
    sys.stderr.write("\n\n-------------------------------------------------\n"
                     "Exception on processing: \n" + msg + "\nAborting!\n"
                     "-------------------------------------------------\n")
    sys.stderr.flush()


def run_backend_without_mask(json_str, lc_field_vals, E_field_vals, *, precision=64, wl_batch=1,
                             use_graph=False, strict=None, stats=None, tuning=None):
    """Propagates the fields of every (wavelength, wavevector, polarisation)
    through the LC sample and applies the screen filter; E_field_vals
    (N_wavelengths, N_wavevectors, 4, Ny, Nx) complex128 is written in place."""
    _check_E(E_field_vals)
    lc = _lib.as_lc(lc_field_vals)
    return _run(json_str, None, lc, None, E_field_vals, precision, wl_batch, use_graph, strict, stats, tuning)


def run_backend_with_mask(json_str, mask_formula, lc_field_vals, mask_vals, E_field_vals, *, precision=64,
                          wl_batch=1, use_graph=False, strict=None, stats=None, tuning=None):
    """As run_backend_without_mask, with a domain mask (mask_vals >= 0 is LC,
    elsewhere the isotropic host) and, if mask_formula is not empty, the
    interface interpolation weights derived from the formula."""
    _check_E(E_field_vals)
    lc = _lib.as_lc(lc_field_vals)
    mask = _lib.as_lc(mask_vals)
    return _run(json_str, mask_formula or "", lc, mask, E_field_vals, precision, wl_batch, use_graph, strict, stats,
                tuning)


def _run(json_str, formula, lc, mask, E, precision, wl_batch, use_graph, strict, stats, tuning=None):
    if use_graph:
        raise ValueError("one-shot backend calls launch eagerly; replay graphs through "
                         "Context(..., use_graph=True), which keeps its buffers between propagations")
    lib = _lib.load()
    opts = _lib.options(precision=precision, wl_batch=wl_batch, use_graph=False, tuning=tuning)
    st = _lib.Stats()
    err = _lib.err_buf()
    Ef = E.reshape(-1).view(np.float64)
    rc = lib.nmk_bpm_run(json_str.encode(), None if formula is None else formula.encode(), _lib.dptr(lc),
                         lc.size, _lib.dptr(mask), 0 if mask is None else mask.size, _lib.dptr(Ef), E.size,
                         ctypes.byref(opts), ctypes.byref(st), err, len(err))
    if stats is not None and rc == 0:
        stats.update(st.as_dict())
    _report(rc, err, strict)
    return None


def propagate(json_str, lc_field_vals, E_shape=None, mask_vals=None, mask_formula="", E_in=None, **kw):
    """Convenience wrapper: allocates and returns the output array."""
    from .fields import output_shape  # noqa: F401  (shape helper lives with the settings builders)
    import json as _json
    if E_shape is None:
        E_shape = output_shape(_json.loads(json_str))
    E = np.zeros(E_shape, dtype=np.complex128) if E_in is None else np.array(E_in, dtype=np.complex128, order="C")
    kw.setdefault("strict", True)
    if mask_vals is None:
        run_backend_without_mask(json_str, lc_field_vals, E, **kw)
    else:
        run_backend_with_mask(json_str, mask_formula, lc_field_vals, mask_vals, E, **kw)
    return E


class Context:
    """Inputs uploaded once to HBM; ``propagate()`` runs the whole step on the
    device and reports device time (used by the benchmark and the serving
    loop)."""

    def __init__(self, json_str, lc_field_vals, n_E, mask_vals=None, mask_formula="", precision=64, wl_batch=1,
                 use_graph=False, profile=False, device=-1, tuning=None):
        self.lib = _lib.load()
        lc = _lib.as_lc(lc_field_vals)
        mask = None if mask_vals is None else _lib.as_lc(mask_vals)
        self.n_E = int(n_E)
        self.json_str = json_str
        opts = _lib.options(precision=precision, wl_batch=wl_batch, use_graph=use_graph, profile=profile,
                            device=device, tuning=tuning)
        err = _lib.err_buf()
        self.h = self.lib.nmk_bpm_create(json_str.encode(), None if mask is None else (mask_formula or "").encode(),
                                         _lib.dptr(lc), lc.size, _lib.dptr(mask), 0 if mask is None else mask.size,
                                         self.n_E, ctypes.byref(opts), err, len(err))
        if not self.h:
            raise BpmBackendError(err.value.decode())
        self.notice = err.value.decode("utf-8", "replace")
        if self.notice:
            sys.stdout.write(self.notice + "\n")
            sys.stdout.flush()

    def upload_fields(self, E):
        _check_E(E)
        err = _lib.err_buf()
        rc = self.lib.nmk_bpm_upload_fields(self.h, _lib.dptr(E.reshape(-1).view(np.float64)), E.size, err, len(err))
        if rc:
            raise BpmBackendError(err.value.decode())

    def propagate(self):
        st = _lib.Stats()
        err = _lib.err_buf()
        rc = self.lib.nmk_bpm_propagate(self.h, ctypes.byref(st), err, len(err))
        if rc:
            raise BpmBackendError(err.value.decode())
        return st.as_dict()

    def set_tuning(self, tuning):
        """Switches the line-solver launch shapes (dict as for ``tuning=``; None: built-in shapes)."""
        opts = _lib.options(tuning=tuning)
        err = _lib.err_buf()
        rc = self.lib.nmk_bpm_set_tuning(self.h, ctypes.byref(opts), err, len(err))
        if rc:
            raise BpmBackendError(err.value.decode())

    def download(self, E):
        _check_E(E)
        err = _lib.err_buf()
        rc = self.lib.nmk_bpm_download(self.h, _lib.dptr(E.reshape(-1).view(np.float64)), E.size, err, len(err))
        if rc:
            raise BpmBackendError(err.value.decode())
        return E

    def close(self):
        if getattr(self, "h", None):
            self.lib.nmk_bpm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
