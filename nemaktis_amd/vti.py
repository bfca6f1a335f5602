"""Reader of the VTK XML image-data files this package writes (the bulk
output of the bpm backend, ``<Results folder name>/<Base name>.vti``): raw
appended little-endian data with UInt64 block headers, uncompressed.  Returns
numpy arrays without needing VTK.

    geo, field_data, point_data = read_vti(path)
    point_data["E_real_inputX_0_0"]      # (Nz, Ny, Nx, 3) float64
"""
import re

import numpy as np

_DTYPES = {"Float64": np.float64, "Float32": np.float32, "Int32": np.int32, "Int64": np.int64,
           "UInt8": np.uint8}


def read_vti(path):
    """Returns (geometry dict, field-data dict, point-data dict) of a raw-appended VTI file."""
    with open(path, "rb") as f:
        raw = f.read()
    start = raw.index(b"<AppendedData")
    mark = raw.index(b"_", start) + 1
    head = raw[:start].decode("ascii")
    if 'header_type="UInt64"' not in head or 'byte_order="LittleEndian"' not in head:
        raise ValueError("only little-endian files with UInt64 headers are read")
    if "compressor=" in head:
        raise ValueError("compressed VTI files are not read")
    img = re.search(r"<ImageData([^>]*)>", head).group(1)
    attr = lambda s, k: re.search(k + r'="([^"]*)"', s).group(1)
    ext = [int(v) for v in attr(img, "WholeExtent").split()]
    geo = {"dims": (ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1),
           "origin": tuple(float(v) for v in attr(img, "Origin").split()),
           "spacing": tuple(float(v) for v in attr(img, "Spacing").split())}
    Nx, Ny, Nz = geo["dims"]

    def arrays(section):
        m = re.search(r"<%s>(.*?)</%s>" % (section, section), head, re.S)
        out = {}
        if not m:
            return out
        for a in re.finditer(r"<DataArray([^>]*)/>", m.group(1)):
            s = a.group(1)
            off = mark + int(attr(s, "offset"))
            n = int(np.frombuffer(raw, np.uint64, 1, off)[0])
            v = np.frombuffer(raw, _DTYPES[attr(s, "type")], n // np.dtype(_DTYPES[attr(s, "type")]).itemsize,
                              off + 8).copy()
            nc = int(re.search(r'NumberOfComponents="(\d+)"', s).group(1)) if "NumberOfComponents" in s else 1
            if section == "PointData":
                v = v.reshape(Nz, Ny, Nx, nc) if nc > 1 else v.reshape(Nz, Ny, Nx)
            out[attr(s, "Name")] = v
        return out

    return geo, arrays("FieldData"), arrays("PointData")


def bulk_fields(path):
    """Bulk output as complex arrays: dict (wavelength, wavevector) -> (2 pols, Nz, Ny, Nx, 2 comps)."""
    geo, fd, pd = read_vti(path)
    out = {}
    for name in pd:
        m = re.match(r"E_real_input([XY])_(\d+)_(\d+)$", name)
        if not m:
            continue
        pol, q, w = "XY".index(m.group(1)), int(m.group(2)), int(m.group(3))
        c = pd[name][..., :2] + 1j * pd[name.replace("E_real_", "E_imag_")][..., :2]
        out.setdefault((w, q), [None, None])[pol] = c
    return {k: np.stack(v) for k, v in out.items()}
