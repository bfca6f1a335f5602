// Settings schema of the bpm-solver backend and the host-side problem
// description handed to the device engine.
//
// The JSON keys, defaults, required set and error texts follow the
// reference's SettingsLib (BeamPropagationSolver/src/SettingsLib/
// RootSettings.cpp:8-12, AlgorithmSettings.cpp:3-22, PhysicsSettings.cpp:8-88,
// PostprocessorSettings.cpp:3-19, BaseSettings.h:22-54) and the checks of
// PythonBackend/bpm_backend.cpp:17-41, 96-121.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "expr_program.h"

namespace nmk {

// error raised for any invalid input; the message mirrors the reference's
struct BpmError : std::runtime_error {
    explicit BpmError(const std::string &m) : std::runtime_error(m) {}
};

enum class BeamKind { Uniform = 0, Gaussian = 1, None = 2 };

struct Settings {
    // Algorithm settings
    std::string lc_field_type, results_folder;
    int n_woodbury = 2, nz_substeps = 1;
    // Physics settings / Initial conditions
    std::string beam_profile, lc_field_file, basis;
    int Nx = 0, Ny = 0, Nz = 0;
    double dx = 0, dy = 0, dz = 0;
    // Physics settings / Coefficients
    std::string no_expr, ne_expr, ne_imag_expr, nhost_expr, nin_expr, nout_expr;
    std::vector<double> wavelengths;
    std::vector<std::pair<double, double>> q;
    // Postprocessor settings
    bool bulk_activate = false, screen_activate = true;
    std::string bulk_name, screen_name;
    std::vector<double> iso_thick, iso_index;
    double z_shift = 0, NA = 0;
};

// Parses and validates the settings JSON (throws BpmError).
Settings parse_settings(const std::string &json_text);

// Everything the device engine needs, with index expressions evaluated.
struct Problem {
    int Nx = 0, Ny = 0, Nz = 0, field_dim = 3;
    double dx = 0, dy = 0, dz = 0;
    std::vector<double> wl, qx, qy;
    // per wavelength
    std::vector<double> no, ne, ne_imag, nhost, nin, nout, nout_fresnel;
    double nout_ref = 1;   // nout(0.6 um), used by the focus correction
    BeamKind beam = BeamKind::Uniform;
    double waist = 0;
    int n_woodbury = 2;
    std::vector<double> iso_thick, iso_index;
    double z_foc = 0, NA = 0;  // z_foc already corrected for the layers
    // bulk output: every transverse plane of the march written as a VTI file
    // (PostprocessorLib/VolumeOutput.cpp:19-32, PostprocessorBase.cpp:24-35)
    bool bulk = false;
    std::string bulk_dir, bulk_path;
    // inputs (borrowed)
    const double *lc = nullptr;
    std::vector<uint8_t> mask;      // empty: every voxel is LC
    std::vector<double> weights;    // empty: every interface weight is 0.5
    // a masked problem validated with its mask preparation deferred: the
    // caller's mask (borrowed) and formula, turned into `mask` / `weights` by
    // prepare_mask
    const double *mask_src = nullptr;
    std::string mask_formula;
    // the interface weights of the mask formula are left to the device engine
    // (weights stays empty): the compiled formula
    bool weights_on_device = false;
    ExprProgram weight_prog;
    int64_t voxels() const { return (int64_t)Nx * Ny * Nz; }
    int64_t n_out_values() const { return (int64_t)wl.size() * (int64_t)qx.size() * 4 * Nx * Ny; }
};

// Builds the problem from the settings and the caller's arrays, applying the
// reference's size checks in the reference's order.  mask == nullptr selects
// the unmasked entry point.
// With defer_mask the checks run (including the formula's syntax) but the
// per-voxel mask and the interface weights are left to prepare_mask, so that
// a caller that queues the problem (the serving loop) does that host work
// later, off the submitting thread.
Problem build_problem(const Settings &s, const double *lc, int64_t n_lc, const double *mask, int64_t n_mask,
                      const char *mask_formula, int64_t n_E, std::string *warning, bool defer_mask = false);
// builds the deferred LC flags of a problem (no-op without) and compiles its
// mask formula for the device; a formula too long for a device program has
// its interface weights computed here on host threads
void prepare_mask(Problem &p, std::string *warning);

// the notice the reference prints when the mask formula's crossings disagree
// with the mask (UtilitiesLib/VectorField.cpp:107-111)
extern const char *const kMaskMismatchNotice;

// Interface interpolation weights of a mask formula (UtilitiesLib/
// VectorField.cpp:64-113) for the mesh and mask of p: w[2*i], w[2*i+1] for
// voxel i of planes 0 .. Nz-2.  When the formula's zero crossings disagree
// with the mask, w is left empty (every weight 0.5) and *warning receives the
// reference's notice.
void interface_weights(const Problem &p, const std::string &formula, std::vector<double> &w, std::string *warning);

}  // namespace nmk
