"""ctypes binding of libnemaktis_bpm.so (the C ABI in include/nemaktis_bpm.h).

The library is built in-tree (``python -m nemaktis_amd.build``); loading
fails loudly when it is missing - there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NEMAKTIS_AMD_LIB selects another build of the library (A/B comparisons)
LIB_PATH = os.environ.get("NEMAKTIS_AMD_LIB") or os.path.join(HERE, "lib", "libnemaktis_bpm.so")

_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("wl_batch", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("device", ctypes.c_int32), ("profile", ctypes.c_int32), ("x_chunks", ctypes.c_int32),
                ("x_rows", ctypes.c_int32), ("y_chunks", ctypes.c_int32), ("y_cols", ctypes.c_int32),
                ("item_order", ctypes.c_int32), ("lean", ctypes.c_int32), ("y_pair", ctypes.c_int32),
                ("x_both", ctypes.c_int32), ("coef_serial", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("setup_ms", ctypes.c_double), ("march_ms", ctypes.c_double), ("output_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("updates", ctypes.c_double), ("updates_per_s", ctypes.c_double),
                ("footprint_bytes", ctypes.c_int64), ("family_ms", ctypes.c_double * 10),
                ("family_launches", ctypes.c_int64 * 10)]

    def as_dict(self):
        names = ["material", "coefs", "boundary", "reserved", "explicit", "solve_x", "solve_y", "output", "filter", "setup"]
        return {
            "setup_ms": self.setup_ms, "march_ms": self.march_ms, "output_ms": self.output_ms,
            "total_ms": self.total_ms, "updates": self.updates, "updates_per_s": self.updates_per_s,
            "footprint_bytes": self.footprint_bytes,
            "family_ms": {n: self.family_ms[i] for i, n in enumerate(names)},
            "family_launches": {n: self.family_launches[i] for i, n in enumerate(names)},
        }


class SrvConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_active", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("max_bypass", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class SrvJobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("t_submit_ms", ctypes.c_double),
                ("t_admit_ms", ctypes.c_double), ("t_end_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("updates", ctypes.c_double), ("footprint_bytes", ctypes.c_int64), ("error", ctypes.c_char * 256),
                ("t_ready_ms", ctypes.c_double), ("notice", ctypes.c_char * 128)]

    STATES = ["queued", "running", "done", "failed", "cancelled"]

    def as_dict(self):
        return {"state": self.STATES[self.state], "status": self.status, "t_submit_ms": self.t_submit_ms,
                "t_admit_ms": self.t_admit_ms, "t_end_ms": self.t_end_ms, "device_ms": self.device_ms,
                "updates": self.updates, "footprint_bytes": self.footprint_bytes,
                "error": self.error.decode("utf-8", "replace"), "t_ready_ms": self.t_ready_ms,
                "notice": self.notice.decode("utf-8", "replace")}


class SrvCounters(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("cancelled", ctypes.c_int64), ("running", ctypes.c_int64), ("queued", ctypes.c_int64),
                ("arena_bytes", ctypes.c_int64), ("arena_used_bytes", ctypes.c_int64),
                ("arena_peak_bytes", ctypes.c_int64), ("peak_active", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("now_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class ViewGeometry(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32), ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("n_wavelengths", ctypes.c_int32), ("wavelengths", c_double_p), ("n_out", c_double_p),
                ("n_wavevectors", ctypes.c_int32), ("wavevectors", c_double_p),
                ("max_NA_objective", ctypes.c_double), ("max_NA_condenser", ctypes.c_double),
                ("N_radial_wavevectors", ctypes.c_int32), ("koehler_1D", ctypes.c_int32),
                ("zfoc_NA_corr", ctypes.c_double), ("device", ctypes.c_int32)]


class ViewOptics(ctypes.Structure):
    _fields_ = [("polariser", ctypes.c_int32), ("analyser", ctypes.c_int32),
                ("polariser_angle", ctypes.c_double), ("analyser_angle", ctypes.c_double),
                ("lower_waveplate", ctypes.c_int32), ("upper_waveplate", ctypes.c_int32),
                ("lower_waveplate_angle", ctypes.c_double), ("upper_waveplate_angle", ctypes.c_double),
                ("intensity", ctypes.c_double), ("NA_condenser", ctypes.c_double),
                ("grayscale", ctypes.c_int32), ("gamma", ctypes.c_int32)]


WAVEPLATES = {"No": 0, "Quarter-wave": 1, "Half-wave": 2, "Tint-sensitive": 3}

EXPORTED = [
    "nmk_bpm_run_backend_without_mask", "nmk_bpm_run_backend_with_mask", "nmk_bpm_default_options",
    "nmk_bpm_run", "nmk_bpm_run_noalloc", "nmk_bpm_create", "nmk_bpm_upload_fields", "nmk_bpm_propagate", "nmk_bpm_download",
    "nmk_bpm_destroy", "nmk_bpm_set_tuning", "nmk_bpm_estimate_footprint", "nmk_bpm_interface_weights",
    "nmk_bpm_interface_weights_device",
    "nmk_bpm_eval_expression", "nmk_bpm_version",
    "nmk_srv_default_config", "nmk_srv_create", "nmk_srv_submit", "nmk_srv_wait", "nmk_srv_cancel",
    "nmk_srv_stats", "nmk_srv_destroy",
    "nmk_view_default_optics", "nmk_view_create", "nmk_view_from_ctx", "nmk_view_set_fields", "nmk_view_focus",
    "nmk_view_get_focused", "nmk_view_render", "nmk_view_get_spectrum", "nmk_view_get_image", "nmk_view_state",
    "nmk_view_timings", "nmk_view_destroy",
]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libnemaktis_bpm.so is not built (run `python -m nemaktis_amd.build`): "
                           "the bpm backend has no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    i64, cp, sz = ctypes.c_int64, ctypes.c_char_p, ctypes.c_size_t
    lib.nmk_bpm_run_backend_without_mask.argtypes = [cp, c_double_p, i64, c_double_p, i64, cp, sz]
    lib.nmk_bpm_run_backend_with_mask.argtypes = [cp, cp, c_double_p, i64, c_double_p, i64, c_double_p, i64, cp, sz]
    lib.nmk_bpm_default_options.argtypes = [ctypes.POINTER(Options)]
    lib.nmk_bpm_run.argtypes = [cp, cp, c_double_p, i64, c_double_p, i64, c_double_p, i64,
                                ctypes.POINTER(Options), ctypes.POINTER(Stats), cp, sz]
    lib.nmk_bpm_run_noalloc.argtypes = [cp, ctypes.c_void_p, i64, ctypes.c_void_p, i64, ctypes.c_void_p, i64,
                                        ctypes.c_void_p, i64, ctypes.c_int32, ctypes.c_void_p, cp, sz]
    lib.nmk_bpm_create.argtypes = [cp, cp, c_double_p, i64, c_double_p, i64, i64, ctypes.POINTER(Options), cp, sz]
    lib.nmk_bpm_create.restype = ctypes.c_void_p
    lib.nmk_bpm_upload_fields.argtypes = [ctypes.c_void_p, c_double_p, i64, cp, sz]
    lib.nmk_bpm_propagate.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats), cp, sz]
    lib.nmk_bpm_download.argtypes = [ctypes.c_void_p, c_double_p, i64, cp, sz]
    lib.nmk_bpm_set_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(Options), cp, sz]
    lib.nmk_bpm_destroy.argtypes = [ctypes.c_void_p]
    lib.nmk_bpm_estimate_footprint.argtypes = [cp, i64, ctypes.c_int, ctypes.POINTER(Options),
                                               ctypes.POINTER(i64), cp, sz]
    lib.nmk_bpm_interface_weights.argtypes = [cp, cp, c_double_p, i64, c_double_p, i64, cp, sz]
    lib.nmk_bpm_interface_weights_device.argtypes = [cp, cp, c_double_p, i64, c_double_p, i64, cp, sz]
    lib.nmk_bpm_eval_expression.argtypes = [cp, ctypes.c_double, c_double_p, cp, sz]
    lib.nmk_bpm_version.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    lib.nmk_srv_default_config.argtypes = [ctypes.POINTER(SrvConfig)]
    lib.nmk_srv_create.argtypes = [ctypes.POINTER(SrvConfig), cp, sz]
    lib.nmk_srv_create.restype = vp
    lib.nmk_srv_submit.argtypes = [vp, cp, cp, c_double_p, i64, c_double_p, i64, c_double_p, i64,
                                   ctypes.POINTER(Options), cp, sz]
    lib.nmk_srv_submit.restype = i64
    lib.nmk_srv_wait.argtypes = [vp, i64, ctypes.c_double, ctypes.POINTER(SrvJobInfo)]
    lib.nmk_srv_cancel.argtypes = [vp, i64]
    lib.nmk_srv_stats.argtypes = [vp, ctypes.POINTER(SrvCounters)]
    lib.nmk_srv_destroy.argtypes = [vp]
    lib.nmk_view_default_optics.argtypes = [ctypes.POINTER(ViewOptics)]
    lib.nmk_view_create.argtypes = [ctypes.POINTER(ViewGeometry), c_double_p, i64, cp, sz]
    lib.nmk_view_create.restype = vp
    lib.nmk_view_from_ctx.argtypes = [vp, ctypes.POINTER(ViewGeometry), cp, sz]
    lib.nmk_view_from_ctx.restype = vp
    lib.nmk_view_set_fields.argtypes = [vp, c_double_p, i64, cp, sz]
    lib.nmk_view_focus.argtypes = [vp, ctypes.c_double, ctypes.c_double, cp, sz]
    lib.nmk_view_get_focused.argtypes = [vp, c_double_p, i64, cp, sz]
    lib.nmk_view_render.argtypes = [vp, ctypes.POINTER(ViewOptics), c_double_p, i64, cp, sz]
    lib.nmk_view_get_spectrum.argtypes = [vp, c_double_p, i64, cp, sz]
    lib.nmk_view_get_image.argtypes = [vp, c_double_p, i64, cp, sz]
    lib.nmk_view_state.argtypes = [vp, c_double_p, c_double_p]
    lib.nmk_view_timings.argtypes = [vp, c_double_p, ctypes.c_int32]
    lib.nmk_view_destroy.argtypes = [vp]
    _lib = lib
    return lib


def dptr(a):
    return None if a is None else a.ctypes.data_as(c_double_p)


def options(precision=64, wl_batch=1, use_graph=False, device=-1, profile=False, tuning=None):
    """tuning: None (built-in shapes) or dict(x_chunks, x_rows, y_chunks, y_cols[, order 0..15][, lean][, y_pair]
    [, x_both][, coef_serial])"""
    o = Options()
    load().nmk_bpm_default_options(ctypes.byref(o))
    o.precision = int(precision)
    o.wl_batch = int(wl_batch)
    o.use_graph = 1 if use_graph else 0
    o.device = int(device)
    o.profile = 1 if profile else 0
    if tuning:
        o.x_chunks, o.x_rows = int(tuning["x_chunks"]), int(tuning["x_rows"])
        o.y_chunks, o.y_cols = int(tuning["y_chunks"]), int(tuning["y_cols"])
        if tuning.get("order") is not None:
            o.item_order = 1 + int(tuning["order"])
        o.lean = 1 if tuning.get("lean") else 0
        o.y_pair = int(tuning.get("y_pair") or 0)
        o.x_both = int(tuning.get("x_both") or 0) & 3
        o.coef_serial = 1 if tuning.get("coef_serial") else 0
    return o


def err_buf():
    return ctypes.create_string_buffer(4096)


def as_lc(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
