"""TEST INFRASTRUCTURE - NOT PRODUCT CODE.

ctypes front-end of the FP64 CPU oracle (``liboracle_bpm.so``, built from
``bpm_oracle.cpp`` by ``oracle/Makefile``).  Only tests/, the smoke check and
bench.py's cpu_baseline leg import this module, as the checker or the CPU
baseline.

``settings_to_params`` turns a bpm-solver settings dictionary (the JSON that
``nemaktis.LightPropagator._bpm_propagation`` sends,
HighLevelPythonInterface/nemaktis/light_propagator.py:152-186) into the
oracle's plain parameter block; index expressions are evaluated here with a
small arithmetic evaluator in ``lambda`` (independent of the product's C++
evaluator, so a parity test also cross-checks the settings parsing).
"""
import ast
import ctypes
import math
import operator
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bpm.so")
_lib = None


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32), ("Nz", ctypes.c_int32),
        ("field_dim", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("n_wl", ctypes.c_int32), ("n_q", ctypes.c_int32),
        ("wl", ctypes.POINTER(ctypes.c_double)), ("q", ctypes.POINTER(ctypes.c_double)),
        ("no", ctypes.POINTER(ctypes.c_double)), ("ne", ctypes.POINTER(ctypes.c_double)),
        ("ne_imag", ctypes.POINTER(ctypes.c_double)), ("nhost", ctypes.POINTER(ctypes.c_double)),
        ("nin", ctypes.POINTER(ctypes.c_double)), ("nout", ctypes.POINTER(ctypes.c_double)),
        ("nout_ref", ctypes.c_double),
        ("beam", ctypes.c_int32), ("waist", ctypes.c_double),
        ("n_woodbury", ctypes.c_int32),
        ("n_iso", ctypes.c_int32),
        ("iso_thick", ctypes.POINTER(ctypes.c_double)), ("iso_index", ctypes.POINTER(ctypes.c_double)),
        ("z_foc", ctypes.c_double), ("NA", ctypes.c_double),
        ("mask", ctypes.POINTER(ctypes.c_double)), ("weights", ctypes.POINTER(ctypes.c_double)),
        ("nthreads", ctypes.c_int32), ("apply_filter", ctypes.c_int32),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        _lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        _lib.oracle_propagate.argtypes = [ctypes.POINTER(OracleParams), P, P]
        _lib.oracle_propagate_bulk.argtypes = [ctypes.POINTER(OracleParams), P, P, P]
        _lib.oracle_create.argtypes = [ctypes.POINTER(OracleParams), P]
        _lib.oracle_create.restype = vp
        _lib.oracle_destroy.argtypes = [vp]
        _lib.oracle_set_wavelength.argtypes = [vp, ctypes.c_int]
        _lib.oracle_nref.argtypes = [vp]
        _lib.oracle_nref.restype = ctypes.c_double
        _lib.oracle_eps_plane.argtypes = [vp, ctypes.c_int, P]
        _lib.oracle_update.argtypes = [vp, ctypes.c_int]
        _lib.oracle_get_R.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        _lib.oracle_block_vmult.argtypes = [vp, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int]
        _lib.oracle_block_solve.argtypes = [vp, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int]
        _lib.oracle_cyclic_line_solve.argtypes = [ctypes.c_int, P, P, P, P, P]
        _lib.oracle_apply_primary.argtypes = [vp, P]
        _lib.oracle_apply_phase.argtypes = [vp, P]
        _lib.oracle_apply_shift.argtypes = [vp, ctypes.c_int, P]
        _lib.oracle_phase_coefs.argtypes = [vp, P]
        _lib.oracle_filter_wavelength.argtypes = [vp, ctypes.c_int, P]
        _lib.oracle_fft2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


# ---------------------------------------------------------------------------
# arithmetic evaluator for index expressions (subset of muParser syntax)
# ---------------------------------------------------------------------------
_BIN = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow}
_FUN = {"sqrt": math.sqrt, "exp": math.exp, "sin": math.sin, "cos": math.cos, "tan": math.tan,
        "log": math.log, "ln": math.log, "log10": math.log10, "abs": abs}


def eval_expr(expr, **variables):
    # muParser's ^ is exponentiation binding tighter than unary minus: Python's **
    node = ast.parse(str(expr).replace("lambda", "lam_").replace("^", "**"), mode="eval").body
    env = {("lam_" if k == "lambda" else k): v for k, v in variables.items()}

    def ev(n):
        if isinstance(n, ast.Constant):
            return float(n.value)
        if isinstance(n, ast.Name):
            if n.id in env:
                return float(env[n.id])
            if n.id == "_pi":
                return math.pi
            raise ValueError("unknown variable " + n.id)
        if isinstance(n, ast.BinOp) and type(n.op) in _BIN:
            return _BIN[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub):
            return -ev(n.operand)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.UAdd):
            return ev(n.operand)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUN:
            return _FUN[n.func.id](*[ev(a) for a in n.args])
        raise ValueError("unsupported expression")
    return ev(node)


def _ring_wavevectors(na, nr):
    """(qx, qy) list of SET/PhysicsSettings.cpp:70-81."""
    q = [(0.0, 0.0)]
    for ir in range(1, nr):
        for iphi in range(6 * ir):
            phi = iphi * 3.1415926535897932 / 3
            q.append((ir * na / (nr - 1) * math.cos(phi), ir * na / (nr - 1) * math.sin(phi)))
    return q


class Params:
    """Keeps the numpy buffers referenced by an OracleParams block alive."""

    def __init__(self, settings, mask=None, weights=None, nthreads=0, apply_filter=True):
        phys = settings["Physics settings"]
        ic = phys["Initial conditions"]
        co = phys["Coefficients"]
        scr = settings["Postprocessor settings"]["Screen output"]
        alg = settings["Algorithm settings"]
        Nx, Ny, Nz = [int(v) for v in ic["Mesh dimensions"]]
        dx, dy, dz = [float(v) for v in ic["Mesh spacings"]]
        if "Wavelengths" in co:
            wl = [float(v) for v in co["Wavelengths"]]
        else:
            mw, fw, nw = float(co["Mean wavelength"]), float(co["Spectral FWHM"]), int(co["N wavelengths"])
            wl = [mw] if nw == 1 else [mw + fw * (i / (nw - 1.0) - 0.5) for i in range(nw)]
        if "Wavevectors" in co:
            raw = [float(v) for v in co["Wavevectors"]]
            q = [(raw[2 * i], raw[2 * i + 1]) for i in range(len(raw) // 2)]
        else:
            q = _ring_wavevectors(float(co["Condenser numerical aperture"]),
                                  int(co["N radial illumination directions"]))
        ev = lambda key, default: [eval_expr(co.get(key, default), **{"lambda": l}) for l in wl]
        self.wl = np.array(wl, dtype=np.float64)
        self.q = np.array(q, dtype=np.float64).reshape(-1)
        self.no = np.array(ev("no", None)); self.ne = np.array(ev("ne", None))
        self.ne_imag = np.array(ev("ne_imag", "0")); self.nhost = np.array(ev("nhost", "1"))
        self.nin = np.array(ev("nin", "1")); self.nout = np.array(ev("nout", "1"))
        self.iso_thick = np.array(scr.get("Isotropic layer thicknesses", []), dtype=np.float64)
        self.iso_index = np.array(scr.get("Isotropic layer refractive indices", []), dtype=np.float64)
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64).reshape(-1)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        beam = ic["Beam profile"]
        waist = 0.0
        if beam == "UniformBeam":
            b = 0
        elif beam == "None":
            b = 2
        elif beam.startswith("GaussianBeam(") and beam.endswith(")"):
            b, waist = 1, float(beam[len("GaussianBeam("):-1])
        else:
            raise ValueError("Unrecognised beam profile")
        fd = 3 if alg["General"]["LC field type"] == "Director" else 6
        self.nwl, self.nq = len(wl), len(q)
        self.shape = (Nx, Ny, Nz)
        self.field_dim = fd
        self.c = OracleParams(
            Nx=Nx, Ny=Ny, Nz=Nz, field_dim=fd, dx=dx, dy=dy, dz=dz,
            n_wl=len(wl), n_q=len(q), wl=_dp(self.wl), q=_dp(self.q),
            no=_dp(self.no), ne=_dp(self.ne), ne_imag=_dp(self.ne_imag), nhost=_dp(self.nhost),
            nin=_dp(self.nin), nout=_dp(self.nout),
            nout_ref=eval_expr(co.get("nout", "1"), **{"lambda": 0.6}),
            beam=b, waist=waist, n_woodbury=int(alg["Beam propagation"]["N Woodbury steps"]),
            n_iso=len(self.iso_thick), iso_thick=_dp(self.iso_thick), iso_index=_dp(self.iso_index),
            z_foc=float(scr["Focalisation z-shift"]), NA=float(scr["Numerical aperture"]),
            mask=_dp(self.mask), weights=_dp(self.weights), nthreads=nthreads,
            apply_filter=1 if apply_filter else 0)


def propagate(settings, lc_vals, E=None, mask=None, weights=None, nthreads=0, apply_filter=True):
    """Runs the oracle; returns E of shape (n_wl, n_q, 4, Ny, Nx) complex128."""
    p = Params(settings, mask, weights, nthreads, apply_filter)
    Nx, Ny, _ = p.shape
    lc = np.ascontiguousarray(lc_vals, dtype=np.float64).reshape(-1)
    if E is None:
        E = np.zeros((p.nwl, p.nq, 4, Ny, Nx), dtype=np.complex128)
    else:
        E = np.ascontiguousarray(E, dtype=np.complex128).copy()
    rc = lib().oracle_propagate(ctypes.byref(p.c), _dp(lc), _dp(E.view(np.float64)))
    if rc != 0:
        raise ValueError("oracle rejected the input (code %d)" % rc)
    return E


def propagate_bulk(settings, lc_vals, E=None, mask=None, weights=None, nthreads=0):
    """Runs the oracle and returns (E, bulk): bulk of shape (n_wl, n_q, 2, Nz, 2, Ny, Nx) complex128 holds
    every transverse plane of the march (BPMIteration.cpp:111-120, 153-168)."""
    p = Params(settings, mask, weights, nthreads, True)
    Nx, Ny, Nz = p.shape
    lc = np.ascontiguousarray(lc_vals, dtype=np.float64).reshape(-1)
    if E is None:
        E = np.zeros((p.nwl, p.nq, 4, Ny, Nx), dtype=np.complex128)
    else:
        E = np.ascontiguousarray(E, dtype=np.complex128).copy()
    bulk = np.zeros((p.nwl, p.nq, 2, Nz, 2, Ny, Nx), dtype=np.complex128)
    rc = lib().oracle_propagate_bulk(ctypes.byref(p.c), _dp(lc), _dp(E.view(np.float64)), _dp(bulk.view(np.float64)))
    if rc != 0:
        raise ValueError("oracle rejected the input (code %d)" % rc)
    return E, bulk


class Session:
    """Operator-level access to one oracle run (one wavelength at a time)."""

    def __init__(self, settings, lc_vals, mask=None, weights=None, nthreads=0):
        self.p = Params(settings, mask, weights, nthreads, False)
        self.lc = np.ascontiguousarray(lc_vals, dtype=np.float64).reshape(-1)
        self.h = lib().oracle_create(ctypes.byref(self.p.c), _dp(self.lc))
        if not self.h:
            raise ValueError("oracle rejected the input")
        self.Nx, self.Ny, self.Nz = self.p.shape
        self.N = self.Nx * self.Ny

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_wavelength(self, iwl):
        lib().oracle_set_wavelength(self.h, iwl)

    @property
    def nref(self):
        return lib().oracle_nref(self.h)

    def eps_plane(self, iz):
        out = np.empty((self.Ny, self.Nx, 12))
        lib().oracle_eps_plane(self.h, iz, _dp(out))
        return out

    def update(self, iz):
        lib().oracle_update(self.h, iz)

    def R(self, t, block, shift):
        out = np.empty((self.Ny, self.Nx), dtype=np.complex128)
        lib().oracle_get_R(self.h, t, block, shift, _dp(out.view(np.float64)))
        return out

    def field(self):
        return np.zeros((2, self.Ny, self.Nx), dtype=np.complex128)

    def block_vmult(self, t, block, dst, dc, src, sc, add):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        lib().oracle_block_vmult(self.h, t, block, _dp(dst.view(np.float64)), dc,
                                 _dp(src.view(np.float64)), sc, 1 if add else 0)

    def block_solve(self, d, block, dst, dc, src, sc):
        src = np.ascontiguousarray(src, dtype=np.complex128)
        lib().oracle_block_solve(self.h, d, block, _dp(dst.view(np.float64)), dc, _dp(src.view(np.float64)), sc)

    def apply_primary(self, f):
        lib().oracle_apply_primary(self.h, _dp(f.view(np.float64)))

    def apply_phase(self, f):
        lib().oracle_apply_phase(self.h, _dp(f.view(np.float64)))

    def apply_shift(self, iq, f):
        lib().oracle_apply_shift(self.h, iq, _dp(f.view(np.float64)))

    def phase_coefs(self):
        out = np.empty((3, self.Ny, self.Nx), dtype=np.complex128)
        lib().oracle_phase_coefs(self.h, _dp(out.view(np.float64)))
        return out

    def filter_wavelength(self, iwl, E_wl):
        lib().oracle_filter_wavelength(self.h, iwl, _dp(E_wl.view(np.float64)))


def cyclic_line_solve(a, b, c, d):
    a, b, c, d = [np.ascontiguousarray(v, dtype=np.complex128) for v in (a, b, c, d)]
    x = np.empty_like(d)
    lib().oracle_cyclic_line_solve(len(d), _dp(a.view(np.float64)), _dp(b.view(np.float64)),
                                   _dp(c.view(np.float64)), _dp(d.view(np.float64)), _dp(x.view(np.float64)))
    return x


def fft2(data, sign):
    data = np.ascontiguousarray(data, dtype=np.complex128).copy()
    Ny, Nx = data.shape[-2:]
    how = int(np.prod(data.shape[:-2])) if data.ndim > 2 else 1
    lib().oracle_fft2(Nx, Ny, how, sign, _dp(data.view(np.float64)))
    return data
