"""Seeded director-field generators and settings builders for the benchmark
workloads of the bpm-solver path.

The layouts follow the reference's conventions: director values are stored as
``(Nz, Ny, Nx, 3)`` float64 (HighLevelPythonInterface/nemaktis/lc_material.py:473-479),
mesh coordinates are centred (BeamPropagationSolver/src/BeamPropagationLib/
BPMIteration.cpp:100-101), and the settings dictionary mirrors what
``LightPropagator._bpm_propagation`` sends
(HighLevelPythonInterface/nemaktis/light_propagator.py:152-186).
"""
import json
import math

import numpy as np


def mesh_coords(Nx, Ny, Nz, dx, dy, dz):
    x = (np.arange(Nx) - (Nx - 1) / 2.0) * dx
    y = (np.arange(Ny) - (Ny - 1) / 2.0) * dy
    z = (np.arange(Nz) - (Nz - 1) / 2.0) * dz
    return x, y, z


def twisted_nematic(Nx, Ny, Nz, total_twist=math.pi / 2):
    """n = (cos phi, sin phi, 0) with phi growing linearly from 0 to total_twist."""
    phi = total_twist * np.arange(Nz) / max(Nz - 1, 1)
    n = np.zeros((Nz, Ny, Nx, 3))
    n[..., 0] = np.cos(phi)[:, None, None]
    n[..., 1] = np.sin(phi)[:, None, None]
    return n


def _gaussian_filter_fft(a, sigma_vox):
    """Periodic Gaussian smoothing of a real 3-D array with per-axis sigma
    (voxels), by a real-to-complex FFT."""
    try:  # multithreaded transforms when scipy is present
        import scipy.fft as sfft
        fwd = lambda x: sfft.rfftn(x, workers=-1)  # noqa: E731
        inv = lambda x, shape: sfft.irfftn(x, s=shape, workers=-1)  # noqa: E731
    except ImportError:
        fwd = np.fft.rfftn
        inv = lambda x, shape: np.fft.irfftn(x, s=shape)  # noqa: E731
    out = fwd(a)
    for ax, s in enumerate(sigma_vox):
        n = a.shape[ax]
        k = (np.fft.rfftfreq(n) if ax == a.ndim - 1 else np.fft.fftfreq(n)) * 2 * math.pi
        g = np.exp(-0.5 * (s * k) ** 2)
        shape = [1] * a.ndim
        shape[ax] = k.size
        out *= g.reshape(shape)
    return inv(out, a.shape)


def random_smooth(Nx, Ny, Nz, spacing, seed, corr_length=1.0):
    """Seeded i.i.d. N(0,1) vector per voxel, Gaussian-smoothed (sigma =
    corr_length, periodic), normalised to unit length."""
    rng = np.random.default_rng(seed)
    sig = [corr_length / spacing[2], corr_length / spacing[1], corr_length / spacing[0]]
    n = np.empty((Nz, Ny, Nx, 3))
    for c in range(3):
        n[..., c] = _gaussian_filter_fft(rng.standard_normal((Nz, Ny, Nx)), sig)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return n


def droplet(Nx, Ny, Nz, spacing, radius):
    """Radial director n = r/|r| (n = z at r = 0) and the mask R^2 - r^2
    (>= 0 inside the droplet), mesh centred on the droplet."""
    x, y, z = mesh_coords(Nx, Ny, Nz, *spacing)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    n = np.zeros((Nz, Ny, Nx, 3))
    safe = np.where(r > 0, r, 1.0)
    n[..., 0] = np.where(r > 0, X / safe, 0.0)
    n[..., 1] = np.where(r > 0, Y / safe, 0.0)
    n[..., 2] = np.where(r > 0, Z / safe, 1.0)
    mask = radius ** 2 - r ** 2
    formula = "%r^2-x^2-y^2-z^2" % radius
    return n, mask, formula


def cholesteric_tilt(Nx, Ny, Nz, spacing, pitch, seed, amplitude=0.2, corr_length=0.5):
    """Cholesteric helix along z (phi = 2 pi z / pitch) with a seeded smooth
    random tilt theta = amplitude * g (g unit-variance Gaussian field)."""
    x, y, z = mesh_coords(Nx, Ny, Nz, *spacing)
    rng = np.random.default_rng(seed)
    sig = [corr_length / spacing[2], corr_length / spacing[1], corr_length / spacing[0]]
    g = _gaussian_filter_fft(rng.standard_normal((Nz, Ny, Nx)), sig)
    g /= max(g.std(), 1e-300)
    th = amplitude * g
    phi = (2 * math.pi * z / pitch)[:, None, None]
    n = np.empty((Nz, Ny, Nx, 3))
    n[..., 0] = np.cos(th) * np.cos(phi)
    n[..., 1] = np.cos(th) * np.sin(phi)
    n[..., 2] = np.sin(th)
    return n


def ring_wavevectors(na, nr):
    """Incidence wavevectors of the condenser rings
    (HighLevelPythonInterface/nemaktis/light_propagator.py:96-103)."""
    q = np.zeros((1 + 3 * nr * (nr - 1), 2))
    for ir in range(1, nr):
        beta = ir * na / (nr - 1)
        for iphi in range(6 * ir):
            phi = iphi * np.pi / (3 * ir)
            q[1 + 3 * ir * (ir - 1) + iphi] = (beta * np.cos(phi), beta * np.sin(phi))
    return q


def make_settings(dims, spacings, wavelengths, wavevectors=((0.0, 0.0),), no="1.5", ne="1.7",
                  ne_imag="0", nhost="1", nin="1", nout="1", na_objective=0.4,
                  beam_profile="UniformBeam", n_woodbury=2, field_type="Director",
                  iso_thicknesses=(), iso_indices=(), z_shift=0.0):
    """Settings dictionary with the keys and nesting the nemaktis caller sends."""
    return {
        "Algorithm settings": {
            "General": {"LC field type": field_type, "Results folder name": ""},
            "Beam propagation": {"N Woodbury steps": int(n_woodbury), "Number of substeps per slab": 1}},
        "Physics settings": {
            "Initial conditions": {
                "Beam profile": beam_profile, "LC field file": "",
                "Mesh dimensions": [int(d) for d in dims],
                "Mesh spacings": [float(s) for s in spacings],
                "Basis convention": "XYZ"},
            "Coefficients": {
                "no": str(no), "ne": str(ne), "ne_imag": str(ne_imag), "nhost": str(nhost),
                "nin": str(nin), "nout": str(nout),
                "Wavelengths": [float(w) for w in wavelengths],
                "Wavevectors": [float(v) for qv in wavevectors for v in qv]}},
        "Postprocessor settings": {
            "Bulk output": {"Activate": False, "Base name": ""},
            "Screen output": {
                "Activate": True, "Base name": "",
                "Isotropic layer thicknesses": [float(t) for t in iso_thicknesses],
                "Isotropic layer refractive indices": [float(n) for n in iso_indices],
                "Focalisation z-shift": float(z_shift),
                "Numerical aperture": float(na_objective)}}}


def settings_json(settings):
    return json.dumps(settings)


def output_shape(settings):
    ic = settings["Physics settings"]["Initial conditions"]
    co = settings["Physics settings"]["Coefficients"]
    Nx, Ny, _ = ic["Mesh dimensions"]
    nwl = len(co["Wavelengths"])
    nq = len(co["Wavevectors"]) // 2
    return (nwl, nq, 4, Ny, Nx)


# ---------------------------------------------------------------------------
# the benchmark workloads (mesh, director and optics of each named case)
# ---------------------------------------------------------------------------
def workload(name, seed=None):
    """Returns (settings, lc_vals (Nz,Ny,Nx,3), mask or None, mask_formula or None)."""
    if name == "tn64":            # 64x64x20 twisted nematic, 1 wavelength, normal incidence
        dims, sp = (64, 64, 20), (0.1, 0.1, 0.1)
        lc = twisted_nematic(64, 64, 20)
        s = make_settings(dims, sp, [0.6], no="1.5", ne="1.7", nhost="1.5", na_objective=0.4)
        return s, lc, None, None
    if name == "random512":       # 512x512x100 smoothed random director
        dims, sp = (512, 512, 100), (0.1, 0.1, 0.1)
        lc = random_smooth(512, 512, 100, sp, seed=2 if seed is None else seed, corr_length=1.0)
        s = make_settings(dims, sp, [0.55], no="1.52", ne="1.74", na_objective=0.4)
        return s, lc, None, None
    if name == "droplet256":      # 256x256x80 radial droplet, 21 wavelengths x 7 wavevectors
        dims, sp = (256, 256, 80), (0.125, 0.125, 0.125)
        lc, mask, formula = droplet(256, 256, 80, sp, 8.0)
        wl = np.linspace(0.40, 0.80, 21)
        s = make_settings(dims, sp, wl, wavevectors=ring_wavevectors(0.2, 2), no="1.5", ne="1.7",
                          nhost="1.51", na_objective=0.4)
        return s, lc, mask, formula
    if name == "cholesteric1024":  # 1024x1024x200 cholesteric + smooth random tilt
        dims, sp = (1024, 1024, 200), (0.05, 0.05, 0.05)
        lc = cholesteric_tilt(1024, 1024, 200, sp, 2.0, seed=4 if seed is None else seed)
        s = make_settings(dims, sp, [0.6], no="1.5", ne="1.7", na_objective=0.4)
        return s, lc, None, None
    raise ValueError("unknown workload " + name)


def updates_per_run(settings):
    """Voxel-slice updates of one run: Nx*Ny*(Nz-1) * n_wavelengths * n_q * 2."""
    ic = settings["Physics settings"]["Initial conditions"]
    co = settings["Physics settings"]["Coefficients"]
    Nx, Ny, Nz = ic["Mesh dimensions"]
    return Nx * Ny * (Nz - 1) * len(co["Wavelengths"]) * (len(co["Wavevectors"]) // 2) * 2
