"""Command line of the bpm-solver on the device:

    python -m nemaktis_amd.cli [ -h | -c FILE | -x FILE ] [--precision 32] [--batch] [--graph]

Mirrors BeamPropagationSolver/src/main.cpp:8-257 and runFromSettings
(SettingsLib/RootSettings.cpp:14-60): -c writes an annotated default settings
file, -x runs a settings file whose lines starting with '#' are comments.  The
director (or Q-tensor) field comes from the "LC field file" VTI
(UtilitiesLib/VTIReader.cpp), the LC domain is where the field is non-zero
(UtilitiesLib/VectorField.cpp:183-195), the bulk output and the screen output
(PostprocessorLib/ScreenOutput.cpp:127-182) are written to the results folder.
The propagation itself runs through the C ABI (nmk_bpm_run); --precision 32,
--batch (all wavelengths at once) and --graph select the device paths.
"""
import json
import os
import re
import sys

import numpy as np

from . import bpm_backend, vti

USAGE = """
Usage: python -m nemaktis_amd.cli [ -h | -c FILE | -x FILE ] [--precision 32] [--batch] [--graph]

\t-h
\t\tPrint this message and exit.

\t-c FILE
\t\tCreate a default settings file with the specified name and exit.

\t-x FILE
\t\tRun the code with the specified settings file.
"""

# the keys and default values of the reference's default settings file
DEFAULT_SETTINGS = """{
\t"Algorithm settings": {
\t\t"General": {
\t\t\t# Field stored in the "LC field file": "Director" (array "n", 3 components) or
\t\t\t# "Q-tensor" (array "Q" with the xx, yy, zz, xy, xz, yz components, or "n" and
\t\t\t# optionally "S", from which Q is built).
\t\t\t"LC field type":\t\t"Director",

\t\t\t# Folder receiving the output files (created when missing).
\t\t\t"Results folder name":\t"results/"
\t\t},
\t\t"Beam propagation": {
\t\t\t# Neumann iterations of each implicit half-step (check convergence by raising it).
\t\t\t"N Woodbury steps": \t\t2,

\t\t\t# Evolution substeps per slab (read, not used).
\t\t\t"Number of substeps per slab": 1
\t\t}
\t},

\t"Physics settings": {
\t\t"Initial conditions": {
\t\t\t# Entrance beam: "UniformBeam" or "GaussianBeam(WAIST)".
\t\t\t"Beam profile":\t\t"UniformBeam",

\t\t\t# VTI file with the LC field (path relative to the working directory).
\t\t\t"LC field file":\t"",

\t\t\t# Axes of the VTI file taken as the solver's X, Y, Z (light travels along Z):
\t\t\t# "XYZ", "YZX" or "ZXY".
\t\t\t"Basis convention":\t"XYZ"
\t\t},

\t\t"Coefficients": {
\t\t\t# Refractive indices; expressions may use the wavelength "lambda" in um.
\t\t\t"no": \t\t\t\t"1.5",
\t\t\t"ne": \t\t\t\t"1.6",
\t\t\t# index outside the LC domain (where the LC field is zero)
\t\t\t"nhost": \t\t\t\t"1.5",
\t\t\t# index of the medium below the sample
\t\t\t"nin": \t\t\t\t"1.5",

\t\t\t# Spectrum of the entrance beam: N wavelengths spread over the FWHM around the mean.
\t\t\t"Mean wavelength":\t0.6,
\t\t\t"Spectral FWHM":\t0.2,
\t\t\t"N wavelengths":\t1,

\t\t\t# Koehler illumination: condenser aperture and Nr rings of incidence directions
\t\t\t# (1 + 3 Nr (Nr - 1) wavevectors).
\t\t\t"Condenser numerical aperture": 0,
\t\t\t"N radial illumination directions": 1
\t\t}
\t},

\t"Postprocessor settings": {
\t\t"Bulk output": {
\t\t\t# Write the field of every transverse plane?
\t\t\t"Activate":\t\t\t\t\t\t\t\tfalse,
\t\t\t"Base name":\t\t\t\t\t\t\t"bulk"
\t\t},

\t\t"Screen output": {
\t\t\t# Write the fields on the screen conjugate to the focal plane?
\t\t\t"Activate":\t\t\t\t\t\t\t\ttrue,
\t\t\t"Base name":\t\t\t\t\t\t\t"screen",

\t\t\t# Isotropic layers between the sample and the objective, in the order light meets them.
\t\t\t"Isotropic layer thicknesses": \t\t\t[1000],
\t\t\t"Isotropic layer refractive indices":\t[1.5],

\t\t\t# Focal plane position relative to the plane conjugate to the sample exit.
\t\t\t"Focalisation z-shift":\t\t\t\t\t0,

\t\t\t# Numerical aperture of the objective.
\t\t\t"Numerical aperture":\t\t\t\t\t0.3
\t\t}
\t}
}
"""


class CliError(Exception):
    pass


def create_default_settings_file(path):
    try:
        with open(path, "w") as f:
            f.write(DEFAULT_SETTINGS)
    except OSError:
        raise CliError("Could not open the specified setting file, maybe you prepended a non-existent directory?")


def minify_json(path):
    """The settings text without empty lines and lines whose first non-blank character is '#'
    (main.cpp:149-167)."""
    try:
        with open(path) as f:
            lines = f.read().split("\n")
    except OSError:
        raise CliError("Could not open the specified setting file")
    keep = [ln for ln in lines if not re.fullmatch(r"\s*", ln) and not re.fullmatch(r"\s*#.*", ln)]
    return "".join(ln + "\n" for ln in keep)


_BASIS = {"XYZ": (0, 1, 2), "YZX": (1, 2, 0), "ZXY": (2, 0, 1)}


def lc_field_from_vti(path, field_type, basis):
    """(lc (Nz, Ny, Nx, 3|6), mask (Nz, Ny, Nx) with +1 inside the LC domain and -1 outside, mesh dimensions,
    mesh spacings) as VTIReader::fill_solution_vector and VectorField::set_mask_from_nonzeros build them."""
    if not path.endswith(".vti"):
        raise CliError("Unrecognised initial solution filetype: must be a path to a vti file.")
    if not os.path.exists(path):
        raise CliError("The given input vti file does not exists")
    geo, _, pd = vti.read_vti(path)
    fd = 3 if field_type == "Director" else 6
    I = _BASIS[basis]
    dim, sp = geo["dims"], geo["spacing"]
    if "n" not in pd and (fd == 3 or "Q" not in pd):
        raise CliError('Could not find the director array "n" in the given vti file')
    # array axes are (z, y, x) = dimension indices (2, 1, 0); output axes follow (I3, I2, I1)
    perm = (2 - I[2], 2 - I[1], 2 - I[0])
    if fd == 3:
        n = pd["n"].reshape(dim[2], dim[1], dim[0], -1).astype(np.float64)
        if np.isnan(n[..., list(I)]).any():
            raise CliError("Error: NaN value in the given vti file")
        lc = np.transpose(n, perm + (3,))[..., list(I)]
    elif "Q" in pd:
        lc = np.transpose(pd["Q"].reshape(dim[2], dim[1], dim[0], 6).astype(np.float64), perm + (3,))
    else:
        n = np.transpose(pd["n"].reshape(dim[2], dim[1], dim[0], -1).astype(np.float64), perm + (3,))
        f = (1 - np.transpose(pd["S"].reshape(dim[2], dim[1], dim[0], -1)[..., 0], perm)) if "S" in pd else 1.0
        lc = np.stack([0.5 * f * (3 * n[..., 0] ** 2 - 1), 0.5 * f * (3 * n[..., 1] ** 2 - 1),
                       0.5 * f * (3 * n[..., 2] ** 2 - 1), 1.5 * f * n[..., 0] * n[..., 1],
                       1.5 * f * n[..., 0] * n[..., 2], 1.5 * f * n[..., 1] * n[..., 2]], axis=-1)
    lc = np.ascontiguousarray(lc)
    mask = np.where(np.sqrt(np.sum(np.abs(lc) ** 2, axis=-1)) > 1e-8, 1.0, -1.0)
    return lc, mask, [dim[i] for i in I], [sp[i] for i in I]


def write_screen_vti(path, settings, E):
    """The screen output file (ScreenOutput.cpp:127-182): one 2-D image, arrays E_real_input{X,Y}_<wl>_<q> and
    E_imag_... with components (Ex, Ey, 0), field data lambda, qx, qy."""
    ic = settings["Physics settings"]["Initial conditions"]
    Nx, Ny, _ = ic["Mesh dimensions"]
    dx, dy, _ = ic["Mesh spacings"]
    nwl, nq = E.shape[:2]
    pd = {}
    for w in range(nwl):
        for q in range(nq):
            for pol, tag in enumerate(("inputX_", "inputY_")):
                for part, fn in (("real", np.real), ("imag", np.imag)):
                    a = np.zeros((Nx * Ny, 3))
                    a[:, 0] = fn(E[w, q, 2 * pol]).reshape(-1)
                    a[:, 1] = fn(E[w, q, 2 * pol + 1]).reshape(-1)
                    pd["E_%s_%s%d_%d" % (part, tag, w, q)] = a
    co = settings["Physics settings"]["Coefficients"]
    wl, qs = _wavelengths_wavevectors(co)
    vti.write_vti(path, (Nx, Ny, 1), (dx, dy, 0.0), (-dx * (Nx - 1) / 2., -dy * (Ny - 1) / 2., 0.0),
                  {"lambda": wl, "qx": [q[0] for q in qs], "qy": [q[1] for q in qs]}, pd)


def _wavelengths_wavevectors(co):
    """The wavelength and wavevector lists the solver derives (SettingsLib/PhysicsSettings.cpp:42-81)."""
    if "Wavelengths" in co:
        wl = [float(v) for v in co["Wavelengths"]]
    else:
        mean, fwhm, n = float(co["Mean wavelength"]), float(co["Spectral FWHM"]), int(co["N wavelengths"])
        wl = [mean] if n == 1 else [mean + fwhm * (i / (n - 1.) - 0.5) for i in range(n)]
    if "Wavevectors" in co:
        v = [float(x) for x in co["Wavevectors"]]
        qs = [(v[2 * i], v[2 * i + 1]) for i in range(len(v) // 2)]
    else:
        na, nr = float(co["Condenser numerical aperture"]), int(co["N radial illumination directions"])
        qs = [(0.0, 0.0)]
        for ir in range(1, nr):
            for iphi in range(6 * ir):
                phi = iphi * np.pi / 3
                qs.append((ir * na / (nr - 1) * np.cos(phi), ir * na / (nr - 1) * np.sin(phi)))
    return wl, qs


def run(settings_path, precision=64, wl_batch=1, use_graph=False):
    """runFromSettings: reads the settings and the LC field file, propagates, writes the outputs; returns
    (settings with the mesh filled in, output fields E)."""
    s = json.loads(minify_json(settings_path))
    gen = s["Algorithm settings"]["General"]
    folder = gen.get("Results folder name", "")
    if folder:
        os.makedirs(folder, exist_ok=True)
    ic = s["Physics settings"]["Initial conditions"]
    lc, mask, dims, spacings = lc_field_from_vti(ic.get("LC field file", ""), gen.get("LC field type", "Director"),
                                                 ic.get("Basis convention", "XYZ"))
    ic["Mesh dimensions"], ic["Mesh spacings"] = dims, spacings
    wl, qs = _wavelengths_wavevectors(s["Physics settings"]["Coefficients"])
    E = np.zeros((len(wl), len(qs), 4, dims[1], dims[0]), np.complex128)
    if use_graph:
        # graph replay lives in a context (one capture, replayed by the march)
        with bpm_backend.Context(json.dumps(s), lc, E.size, mask_vals=mask, mask_formula="", precision=precision,
                                 wl_batch=wl_batch, use_graph=True) as ctx:
            ctx.propagate()
            ctx.download(E)
    else:
        bpm_backend.run_backend_with_mask(json.dumps(s), "", lc, mask, E, precision=precision, wl_batch=wl_batch,
                                          strict=True)
    scr = s["Postprocessor settings"]["Screen output"]
    if scr.get("Activate", True):
        base = scr.get("Base name", "")
        path = (folder.rstrip("/") + "/" if folder not in ("", ".") else "") + base + ".vti"
        write_screen_vti(path, s, E)
    return s, E


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    opts = {"precision": 64, "wl_batch": 1, "use_graph": False}
    rest = []
    i = 0
    while i < len(argv):
        a = argv[i]
        if a == "--precision" and i + 1 < len(argv):
            opts["precision"] = int(argv[i + 1])
            i += 2
            continue
        if a == "--batch":
            opts["wl_batch"] = 0
        elif a == "--graph":
            opts["use_graph"] = True
        else:
            rest.append(a)
        i += 1
    if not rest:
        print("\nSyntax error!")
        print(USAGE)
        return 0
    try:
        if rest[0] == "-h":
            print(USAGE)
            return 0
        if rest[0] == "-c" and len(rest) > 1 and not rest[1].startswith("-"):
            create_default_settings_file(rest[1])
            return 0
        if rest[0] == "-x" and len(rest) > 1 and not rest[1].startswith("-"):
            run(rest[1], **opts)
            return 0
        print("\nSyntax error!")
        print(USAGE)
        return 0
    except (CliError, bpm_backend.BpmBackendError, KeyError, ValueError) as e:
        msg = str(e) if not isinstance(e, KeyError) else "missing setting %s" % e
        print("\n\n-------------------------------------------------\nException on processing: \n%s\nAborting!\n"
              "-------------------------------------------------" % msg, file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
