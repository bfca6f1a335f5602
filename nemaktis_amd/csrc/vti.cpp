#include "vti.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <thread>
#include <vector>
#include <algorithm>

#include "settings.h"

namespace nmk {

namespace {

std::string num(double v) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

void put(FILE *f, const void *p, size_t n, const std::string &path) {
    if (n && std::fwrite(p, 1, n, f) != n) throw BpmError("cannot write the bulk output file " + path);
}

}  // namespace

void write_bulk_vti(const std::string &path, const std::string &dir, const BulkGeometry &g,
                    const std::function<void(size_t, c64 *)> &fetch) {
    if (!dir.empty() && ::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST)
        throw BpmError("cannot create the results folder " + dir + ": " + std::strerror(errno));
    const size_t N = (size_t)g.Nx * g.Ny, npts = N * g.Nz;
    const size_t nblocks = g.wl.size() * g.qx.size() * 2;
    const uint64_t arr_bytes = (uint64_t)npts * 3 * sizeof(double);

    // header: field data first, then the point-data arrays in the order the
    // reference adds them (wavelength, wavevector, polarisation; real, imag)
    std::ostringstream h;
    uint64_t off = 0;
    auto field = [&](const char *name, size_t n) {
        h << "      <DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfTuples=\"" << n
          << "\" format=\"appended\" offset=\"" << off << "\"/>\n";
        off += 8 + 8 * (uint64_t)n;
    };
    const std::string ext = "0 " + std::to_string(g.Nx - 1) + " 0 " + std::to_string(g.Ny - 1) + " 0 " +
                            std::to_string(g.Nz - 1);
    h << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
      << "  <ImageData WholeExtent=\"" << ext << "\" Origin=\"" << num(-g.dx * (g.Nx - 1) / 2.) << " "
      << num(-g.dy * (g.Ny - 1) / 2.) << " " << num(-g.dz * (g.Nz - 1) / 2.) << "\" Spacing=\"" << num(g.dx) << " "
      << num(g.dy) << " " << num(g.dz) << "\" Direction=\"1 0 0 0 1 0 0 0 1\">\n"
      << "    <FieldData>\n";
    field("lambda", g.wl.size());
    field("qx", g.qx.size());
    field("qy", g.qy.size());
    h << "    </FieldData>\n"
      << "    <Piece Extent=\"" << ext << "\">\n"
      << "      <PointData>\n";
    const char *pol_str[2] = {"inputX_", "inputY_"};
    for (size_t w = 0; w < g.wl.size(); w++)
        for (size_t q = 0; q < g.qx.size(); q++)
            for (int pol = 0; pol < 2; pol++)
                for (const char *part : {"real", "imag"}) {
                    h << "        <DataArray type=\"Float64\" Name=\"E_" << part << "_" << pol_str[pol] << q << "_"
                      << w << "\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << off << "\"/>\n";
                    off += 8 + arr_bytes;
                }
    h << "      </PointData>\n"
      << "      <CellData>\n"
      << "      </CellData>\n"
      << "    </Piece>\n"
      << "  </ImageData>\n"
      << "  <AppendedData encoding=\"raw\">\n"
      << "   _";
    const std::string head = h.str();

    File out;
    out.f = std::fopen(path.c_str(), "wb");
    if (!out.f) throw BpmError("cannot open the bulk output file " + path + ": " + std::strerror(errno));
    put(out.f, head.data(), head.size(), path);
    auto vec = [&](const std::vector<double> &v) {
        const uint64_t n = 8 * (uint64_t)v.size();
        put(out.f, &n, 8, path);
        put(out.f, v.data(), n, path);
    };
    vec(g.wl);
    vec(g.qx);
    vec(g.qy);
    std::vector<c64> blk(npts * 2);
    std::vector<double> re(npts * 3), im(npts * 3);
    // planes split over host threads: (Ex, Ey) complex planes -> 3-component real and imaginary arrays
    const int nthreads = (int)std::max<size_t>(1, std::min<size_t>({8, std::max(1u, std::thread::hardware_concurrency()),
                                                                     (size_t)g.Nz, npts >> 18}));
    auto planes = [&](int z0, int z1) {
        for (int z = z0; z < z1; z++) {
            const c64 *ex = blk.data() + (size_t)z * 2 * N, *ey = ex + N;
            double *r = re.data() + (size_t)z * N * 3, *i = im.data() + (size_t)z * N * 3;
            for (size_t m = 0; m < N; m++) {
                r[3 * m] = ex[m].x;
                r[3 * m + 1] = ey[m].x;
                r[3 * m + 2] = 0.0;
                i[3 * m] = ex[m].y;
                i[3 * m + 1] = ey[m].y;
                i[3 * m + 2] = 0.0;
            }
        }
    };
    for (size_t k = 0; k < nblocks; k++) {
        fetch(k, blk.data());
        if (nthreads == 1) {
            planes(0, g.Nz);
        } else {
            std::vector<std::thread> th;
            for (int t = 1; t < nthreads; t++) th.emplace_back(planes, g.Nz * t / nthreads, g.Nz * (t + 1) / nthreads);
            planes(0, g.Nz / nthreads);
            for (auto &x : th) x.join();
        }
        put(out.f, &arr_bytes, 8, path);
        put(out.f, re.data(), arr_bytes, path);
        put(out.f, &arr_bytes, 8, path);
        put(out.f, im.data(), arr_bytes, path);
    }
    const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
    put(out.f, tail, sizeof(tail) - 1, path);
    if (std::fclose(out.f) != 0) {
        out.f = nullptr;
        throw BpmError("cannot write the bulk output file " + path);
    }
    out.f = nullptr;
}

}  // namespace nmk
