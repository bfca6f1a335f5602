// VTK XML image-data writer of the bulk (per-plane) optical fields, the file
// the reference's VolumeOutput writes with vtkXMLImageDataWriter
// (PostprocessorLib/VolumeOutput.cpp:19-32) from the arrays laid out by
// BulkOpticalFieldCollection (BeamPropagationLib/OpticalFieldCollection.cpp:75-148).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "device_common.h"

namespace nmk {

struct BulkGeometry {
    int Nx = 0, Ny = 0, Nz = 0;
    double dx = 0, dy = 0, dz = 0;
    std::vector<double> wl, qx, qy;
};

// Writes `path` (creating `dir` first when it is not empty).  Block k =
// pol + 2 * (q + Nq * wavelength) holds Nz * 2 * Nx * Ny complex values in
// the order [plane][component][point]; fetch(k, dst) copies it to dst.  The
// point data are the arrays E_real_input{X,Y}_<q>_<wavelength> and
// E_imag_..., three components (Ex, Ey, 0) per point, point index
// ix + Nx * (iy + Ny * iz); the field data "lambda", "qx", "qy"; appended raw
// little-endian data with UInt64 block headers.
void write_bulk_vti(const std::string &path, const std::string &dir, const BulkGeometry &g,
                    const std::function<void(size_t, c64 *)> &fetch);

}  // namespace nmk
