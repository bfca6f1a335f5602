// Microscope imaging of the output fields on the device: the objective
// refocus of OpticalFields.focus_fields
// (HighLevelPythonInterface/nemaktis/light_propagator.py:578-623, 695-719)
// and the micrograph synthesis of FieldViewer.update_image
// (HighLevelPythonInterface/nemaktis/field_viewer.py:340-470, 472-517).
//
// Data layout in HBM (caller layout of the reference, complex128):
//   fields, focused  E[((l*Nq + q)*4 + c)*N + m]    c: Ex, Ey (input x), Ex, Ey (input y)
//   filter           F[(l*Nq + q)*N + m]             objective mask x defocus phase / N
//   spectrum         S[l*N + m] on the device, (Ny, Nx, Nl) when downloaded
//   image            I[m*3 + channel]
// with m = ix + Nx*iy and N = Nx*Ny.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_common.h"
#include "kernels.h"

namespace nmk {

struct ViewGeometry {
    int Nx = 0, Ny = 0;
    double Lx = 0, Ly = 0;
    std::vector<double> wl, n_out;  // per wavelength
    std::vector<double> qx, qy;     // per wavevector
    double max_NA_objective = 0, max_NA_condenser = 0;
    int Nr = 1;
    bool koehler_1D = false;
    double zfoc_NA_corr = 0;
    int device = -1;
    int64_t n_values() const { return (int64_t)wl.size() * (int64_t)qx.size() * 4 * Nx * Ny; }
};

// wavevectors of the condenser rings (light_propagator.py:578-599)
void condenser_wavevectors(ViewGeometry &g);
// checks sizes, fills defaults (throws BpmError)
void validate_geometry(ViewGeometry &g);

enum Waveplate { WP_NONE = 0, WP_QUARTER = 1, WP_HALF = 2, WP_TINT = 3 };

struct ViewOptics {
    bool polariser = true, analyser = true;
    double polariser_angle = 0, analyser_angle = 90;
    int lower_waveplate = WP_NONE, upper_waveplate = WP_NONE;
    double lower_waveplate_angle = 0, upper_waveplate_angle = 0;
    double intensity = 1, NA_condenser = 0;
    bool grayscale = false, gamma = true;
};

// q-integration of FieldViewer._update_averaged_specter: per-wavevector
// factors (zero outside the used range, the 2/NA^2 or 1/(2 NA) prefactor
// folded in) and the range [q0, q1) of FieldViewer._update_specter
struct QuadRule {
    std::vector<double> w;
    int q0 = 0, q1 = 0;
};
QuadRule condenser_quadrature(const ViewGeometry &g, double NA_condenser);

// per-wavelength Jones matrices: T = A (Wu (S R)) with R = Wl P
struct JonesSet {
    c64 R[4], Wu[4], A[4];
};
std::vector<JonesSet> jones_matrices(const ViewGeometry &g, const ViewOptics &o);

// per (wavelength, wavevector) scalars of the objective filter
struct ViewPlane {
    double k0, kn, px, py;
};

// kernel launchers (viewer_kernels.hip)
void launch_view_filter(c64 *filt, int Nx, int Ny, double dx, double dy, const ViewPlane *planes, int nplanes,
                        double NA, double zf, hipStream_t st);
void launch_view_columns(c64 *data, const c64 *filt, int Nx, int Ny, int nlq, const c64 *tw_y, const FftPlan &py,
                         hipStream_t st);
void launch_view_spectrum(const c64 *E, int64_t N, int Nl, int Nq, int q0, int q1, const double *w,
                          const JonesSet *jones, double *spec, hipStream_t st);
void launch_view_colour(const double *spec, int64_t N, int Nl, const double *cmf, double intensity, int grayscale,
                        int gamma, double *spec_out, double *image, hipStream_t st);

// sRGB transfer function with clipping to [0, 1] (host copy, for tests)
double srgb_gamma(double v);

class Viewer {
public:
    explicit Viewer(const ViewGeometry &g);
    ~Viewer();
    Viewer(const Viewer &) = delete;
    Viewer &operator=(const Viewer &) = delete;

    const ViewGeometry &geometry() const { return g_; }
    void set_fields(const c64 *E_host);
    // device-to-device copy of a caller's resident fields (same layout)
    void set_fields_device(const c64 *E_dev, hipStream_t producer);
    // OpticalFields.update_NA_objective (when NA is not NaN) + focus_fields
    // (z kept when NaN)
    void focus(double z_focus, double NA_objective);
    void download_focused(c64 *E_host);
    // FieldViewer._update_specter + _update_averaged_specter, and the colour
    // conversion when cmf (Nl x 3) is given
    void render(const ViewOptics &o, const double *cmf);
    void download_spectrum(double *out);  // (Ny, Nx, Nl)
    void download_image(double *out);     // (Ny, Nx, 3)
    double z_focus() const { return z_; }
    double NA_objective() const { return NA_; }
    bool has_image() const { return has_image_; }
    // device ms of the last calls: filter, rows forward, columns, rows
    // backward, spectrum, colour
    enum { T_FILTER = 0, T_ROWS_FWD, T_COLUMNS, T_ROWS_BWD, T_SPECTRUM, T_COLOUR, T_COUNT };
    const double *timings() const { return ms_; }
    hipStream_t stream() const { return st_; }

private:
    ViewGeometry g_;
    int Nl_, Nq_;
    int64_t N_;
    double dx_, dy_;
    double z_ = 0, NA_ = 0;
    bool focused_ = false, has_spec_ = false, has_image_ = false;
    hipStream_t st_ = nullptr;
    hipEvent_t ev_[7] = {};
    FftPlan px_{}, py_{};
    c64 *vals_ = nullptr, *focused_buf_ = nullptr, *filt_ = nullptr, *twx_ = nullptr, *twy_ = nullptr;
    ViewPlane *planes_ = nullptr;
    double *spec_ = nullptr, *spec_out_ = nullptr, *image_ = nullptr, *w_ = nullptr, *cmf_ = nullptr;
    JonesSet *jones_ = nullptr;
    std::vector<double> h_w_, h_cmf_;
    std::vector<JonesSet> h_jones_;
    double ms_[T_COUNT] = {0};
    void release();
};

}  // namespace nmk
