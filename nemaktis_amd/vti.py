"""VTK XML image-data files (.vti) without VTK: the director input of the
bpm-solver command line (UtilitiesLib/VTIReader.cpp reads them with
vtkXMLImageDataReader), and the screen and bulk outputs the solver writes
(PostprocessorLib/ScreenOutput.cpp:127-182, VolumeOutput.cpp:19-32).

Reading covers what VTK writes for image data: inline "ascii" and "binary"
(base64) arrays, "appended" data in raw or base64 encoding, UInt32 or UInt64
block headers, either byte order, and vtkZLibDataCompressor blocks.  Writing
produces uncompressed raw appended little-endian data with UInt64 headers.

    geo, field_data, point_data = read_vti(path)
    point_data["E_real_inputX_0_0"]      # (Nz, Ny, Nx, 3) float64
"""
import base64
import re
import zlib

import numpy as np

_TYPES = {"Float64": "f8", "Float32": "f4", "Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2",
          "Int32": "i4", "UInt32": "u4", "Int64": "i8", "UInt64": "u8"}


def _attr(s, k, default=None):
    m = re.search(r'\b' + k + r'="([^"]*)"', s)
    return m.group(1) if m else default


class _Decoder:
    """Block decoding of one file (header type, byte order, compressor)."""

    def __init__(self, head):
        vtk = re.search(r"<VTKFile([^>]*)>", head).group(1)
        self.order = ">" if _attr(vtk, "byte_order", "LittleEndian") == "BigEndian" else "<"
        self.htype = np.dtype(self.order + ("u8" if _attr(vtk, "header_type", "UInt32") == "UInt64" else "u4"))
        comp = _attr(vtk, "compressor")
        if comp not in (None, "vtkZLibDataCompressor"):
            raise ValueError("unsupported VTK compressor " + comp)
        self.zlib = comp is not None

    def raw_block(self, buf, pos):
        """Decodes the binary block at buf[pos:] (raw bytes); returns (payload, bytes consumed)."""
        h = self.htype.itemsize
        if not self.zlib:
            n = int(np.frombuffer(buf, self.htype, 1, pos)[0])
            return bytes(buf[pos + h:pos + h + n]), h + n
        nb = int(np.frombuffer(buf, self.htype, 1, pos)[0])
        sizes = np.frombuffer(buf, self.htype, 3 + nb, pos).astype(np.int64)
        start = pos + h * (3 + nb)
        out = []
        for c in sizes[3:]:
            out.append(zlib.decompress(bytes(buf[start:start + c])))
            start += c
        return b"".join(out), start - pos

    def b64_block(self, text):
        """Decodes a base64 block (inline binary or base64-appended data)."""
        h = self.htype.itemsize
        if not self.zlib:
            data = base64.b64decode(text)
            n = int(np.frombuffer(data, self.htype, 1, 0)[0])
            return data[h:h + n]
        # the header and the compressed blocks are encoded separately
        first = base64.b64decode(text[:((h + 3 - 1) // 3) * 4 + 4])
        nb = int(np.frombuffer(first, self.htype, 1, 0)[0])
        hlen = ((h * (3 + nb) + 2) // 3) * 4
        sizes = np.frombuffer(base64.b64decode(text[:hlen]), self.htype, 3 + nb, 0).astype(np.int64)
        payload = base64.b64decode(text[hlen:])
        out, start = [], 0
        for c in sizes[3:]:
            out.append(zlib.decompress(payload[start:start + c]))
            start += c
        return b"".join(out)


def read_vti(path):
    """Returns (geometry dict, field-data dict, point-data dict) of a .vti file; point arrays are shaped
    (Nz, Ny, Nx) or (Nz, Ny, Nx, components)."""
    with open(path, "rb") as f:
        raw = f.read()
    app = raw.find(b"<AppendedData")
    head = raw[:app if app >= 0 else len(raw)].decode("latin-1")
    dec = _Decoder(head)
    mark, app_end, app_b64 = None, None, False
    if app >= 0:
        tag_end = raw.index(b">", app)
        app_b64 = _attr(raw[app:tag_end].decode("latin-1"), "encoding", "raw") == "base64"
        mark = raw.index(b"_", tag_end) + 1
        app_end = raw.rfind(b"</AppendedData>")
    img = re.search(r"<ImageData([^>]*)>", head).group(1)
    ext = [int(v) for v in _attr(img, "WholeExtent").split()]
    geo = {"dims": (ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1),
           "origin": tuple(float(v) for v in _attr(img, "Origin", "0 0 0").split()),
           "spacing": tuple(float(v) for v in _attr(img, "Spacing", "1 1 1").split())}
    Nx, Ny, Nz = geo["dims"]
    # base64 appended arrays are concatenated: each runs to the next array's offset
    offsets = sorted(set(int(o) for o in re.findall(r'offset="(\d+)"', head)))

    def values(attrs, body):
        dt = np.dtype(dec.order + _TYPES[_attr(attrs, "type")])
        fmt = _attr(attrs, "format", "ascii")
        if fmt == "ascii":
            return np.array(body.split(), dtype=dt.newbyteorder("="))
        if fmt == "binary":
            data = dec.b64_block("".join(body.split()))
        elif app_b64:
            off = int(_attr(attrs, "offset"))
            nxt = [o for o in offsets if o > off]
            stop = mark + nxt[0] if nxt else app_end
            data = dec.b64_block("".join(raw[mark + off:stop].decode("ascii").split()))
        else:
            data, _ = dec.raw_block(raw, mark + int(_attr(attrs, "offset")))
        return np.frombuffer(data, dt).astype(dt.newbyteorder("="))

    def arrays(section):
        m = re.search(r"<%s[^>]*>(.*?)</%s>" % (section, section), head, re.S)
        out = {}
        if not m:
            return out
        for a in re.finditer(r"<DataArray([^>]*?)(/>|>(.*?)</DataArray>)", m.group(1), re.S):
            attrs, body = a.group(1), a.group(3) or ""
            v = values(attrs, body)
            nc = int(_attr(attrs, "NumberOfComponents", "1"))
            if section == "PointData":
                v = v.reshape(Nz, Ny, Nx, nc) if nc > 1 else v.reshape(Nz, Ny, Nx)
            out[_attr(attrs, "Name")] = v
        return out

    return geo, arrays("FieldData"), arrays("PointData")


def write_vti(path, dims, spacing, origin, field_data, point_data):
    """Writes image data with raw appended little-endian arrays and UInt64 headers.  field_data: ordered
    {name: 1-D float64}; point_data: ordered {name: (Nx*Ny*Nz, components) or (Nx*Ny*Nz,) float64}, point
    index ix + Nx*(iy + Ny*iz)."""
    Nx, Ny, Nz = dims
    ext = "0 %d 0 %d 0 %d" % (Nx - 1, Ny - 1, Nz - 1)
    num = lambda v: repr(float(v))
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
             '  <ImageData WholeExtent="%s" Origin="%s" Spacing="%s" Direction="1 0 0 0 1 0 0 0 1">'
             % (ext, " ".join(map(num, origin)), " ".join(map(num, spacing))), "    <FieldData>"]
    blobs, off = [], 0

    def add(v):
        nonlocal off
        b = np.ascontiguousarray(v, dtype="<f8").tobytes()
        blobs.append(np.uint64(len(b)).astype("<u8").tobytes() + b)
        o = off
        off += 8 + len(b)
        return o

    for name, v in field_data.items():
        v = np.atleast_1d(np.asarray(v, dtype=np.float64))
        lines.append('      <DataArray type="Float64" Name="%s" NumberOfTuples="%d" format="appended" offset="%d"/>'
                     % (name, v.size, add(v)))
    lines += ["    </FieldData>", '    <Piece Extent="%s">' % ext, "      <PointData>"]
    for name, v in point_data.items():
        v = np.asarray(v, dtype=np.float64)
        nc = 1 if v.ndim == 1 else v.shape[1]
        if v.shape[0] != Nx * Ny * Nz:
            raise ValueError("point array %s has %d tuples for %d points" % (name, v.shape[0], Nx * Ny * Nz))
        lines.append('        <DataArray type="Float64" Name="%s" NumberOfComponents="%d" format="appended" '
                     'offset="%d"/>' % (name, nc, add(v)))
    lines += ["      </PointData>", "      <CellData>", "      </CellData>", "    </Piece>", "  </ImageData>",
              '  <AppendedData encoding="raw">', "   _"]
    with open(path, "wb") as f:
        f.write("\n".join(lines).encode("ascii"))
        for b in blobs:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


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
