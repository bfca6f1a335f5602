/*
 * nemaktis_view.h - C ABI of the microscope imaging step on the device:
 * objective refocus of the output fields and micrograph synthesis, the
 * steps Nemaktis runs on the host after the bpm backend.
 *
 * Replaces, for the Python caller:
 *   OpticalFields (HighLevelPythonInterface/nemaktis/light_propagator.py):
 *     constructor, manual form (:561-623), update_NA_objective (:690-693),
 *     focus_fields (:696-719), focused_vals (:649-653)
 *   FieldViewer (HighLevelPythonInterface/nemaktis/field_viewer.py):
 *     update_image (:340-344) = _update_specter (:347-385),
 *     _update_averaged_specter (:388-446), _update_POM_image (:449-455),
 *     get_image (:330-333), get_spectrum (:335-338), Jones matrices (:472-512)
 *
 * Arrays (same conventions as nemaktis_bpm.h):
 *   fields     (N_wavelengths, N_wavevectors, 4, Ny, Nx) complex128 as
 *              interleaved (re, im) float64, n_E counts complex values
 *   spectrum   (Ny, Nx, N_wavelengths) float64, q-averaged, before intensity
 *   image      (Ny, Nx, 3) float64 in [0, 1]
 *   cmf        (N_wavelengths, 3) float64 colour matching table (X, Y, Z per
 *              wavelength, the `cmf` argument of FieldViewer)
 * Rows come in mesh order (iy = 0 first); FieldViewer.get_image and
 * get_spectrum flip them and tile the micrograph on the host.  Every entry
 * point returns an nmk_bpm_status code with the message in err_msg.
 */
#ifndef NEMAKTIS_VIEW_H
#define NEMAKTIS_VIEW_H

#include "nemaktis_bpm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nmk_view nmk_view;

enum nmk_view_waveplate {
    NMK_VIEW_NO_WAVEPLATE = 0,
    NMK_VIEW_QUARTER_WAVE = 1,   /* "Quarter-wave" (achromatic) */
    NMK_VIEW_HALF_WAVE = 2,      /* "Half-wave" (achromatic) */
    NMK_VIEW_TINT_SENSITIVE = 3  /* "Tint-sensitive" (full wave at 540 nm) */
};

/* the keyword arguments of the OpticalFields constructor */
typedef struct {
    int32_t Nx, Ny;               /* mesh_dimensions */
    double Lx, Ly;                /* mesh_lengths (spacing L / max(1, N - 1)) */
    int32_t n_wavelengths;
    const double *wavelengths;    /* um */
    const double *n_out;          /* per wavelength; NULL: 1 */
    int32_t n_wavevectors;        /* 0: the condenser rings of N_radial_wavevectors */
    const double *wavevectors;    /* (qx, qy) pairs */
    double max_NA_objective, max_NA_condenser;
    int32_t N_radial_wavevectors;
    int32_t koehler_1D;
    double zfoc_NA_corr;
    int32_t device;               /* HIP device, -1: current */
} nmk_view_geometry;

/* the optical elements of FieldViewer (class attributes) */
typedef struct {
    int32_t polariser, analyser;                /* 0 / 1 */
    double polariser_angle, analyser_angle;     /* degrees */
    int32_t lower_waveplate, upper_waveplate;   /* nmk_view_waveplate */
    double lower_waveplate_angle, upper_waveplate_angle;
    double intensity;
    double NA_condenser;
    int32_t grayscale;
    int32_t gamma;                              /* 1: sRGB transfer function, 0: clip only */
} nmk_view_optics;

/* FieldViewer's defaults: polariser 0 deg, analyser 90 deg, no waveplates,
 * intensity 1, NA_condenser 0, colour, gamma */
void nmk_view_default_optics(nmk_view_optics *o);

/* Device copy of a set of fields (E may be NULL: zeros).  No counterpart
 * besides the OpticalFields constructor: the fields live in HBM from here on. */
nmk_view *nmk_view_create(const nmk_view_geometry *geometry, const double *E, int64_t n_E, char *err_msg,
                          size_t err_len);

/* Fields of the last nmk_bpm_propagate of ctx, copied on the device (no host
 * round trip).  geometry may be NULL: mesh, wavelengths, n_out, wavevectors,
 * objective NA and focus correction are then taken from the ctx's settings as
 * LightPropagator._bpm_propagation builds them (light_propagator.py:196-219),
 * and N_radial_wavevectors / max_NA_condenser from the wavevectors as the VTI
 * import does (light_propagator.py:497-507). */
nmk_view *nmk_view_from_ctx(nmk_bpm_ctx *ctx, const nmk_view_geometry *geometry, char *err_msg, size_t err_len);

/* OpticalFields.vals setter (light_propagator.py:670-675) */
int nmk_view_set_fields(nmk_view *v, const double *E, int64_t n_E, char *err_msg, size_t err_len);

/* update_NA_objective(NA_objective) unless NaN, then focus_fields(z_focus),
 * z kept when NaN (light_propagator.py:690-719) */
int nmk_view_focus(nmk_view *v, double z_focus, double NA_objective, char *err_msg, size_t err_len);

/* OpticalFields.focused_vals */
int nmk_view_get_focused(nmk_view *v, double *E, int64_t n_E, char *err_msg, size_t err_len);

/* FieldViewer.update_image: spectrum always, image when cmf is given
 * (n_cmf = 3 * N_wavelengths) */
int nmk_view_render(nmk_view *v, const nmk_view_optics *optics, const double *cmf, int64_t n_cmf, char *err_msg,
                    size_t err_len);

/* results of the last render (n = Ny * Nx * N_wavelengths, n = Ny * Nx * 3) */
int nmk_view_get_spectrum(nmk_view *v, double *spectrum, int64_t n, char *err_msg, size_t err_len);
int nmk_view_get_image(nmk_view *v, double *image, int64_t n, char *err_msg, size_t err_len);

/* current z_focus and objective NA (OpticalFields.z_focus, FieldViewer.NA_objective) */
int nmk_view_state(const nmk_view *v, double *z_focus, double *NA_objective);

/* device ms of the last focus / render: [0] filter, [1] rows forward,
 * [2] columns (forward, filter, backward), [3] rows backward, [4] spectrum,
 * [5] colour; returns the number of values written */
int nmk_view_timings(const nmk_view *v, double *ms, int32_t n);

void nmk_view_destroy(nmk_view *v);

#ifdef __cplusplus
}
#endif

#endif /* NEMAKTIS_VIEW_H */
