/*
 * TEST INFRASTRUCTURE - NOT PRODUCT CODE.
 *
 * FP64 CPU restatement of the Nemaktis bpm-solver propagation step (the
 * BPMIteration z-march with its permittivity, phase, shift, paraxial ADI,
 * Fresnel and screen-filter stages).  It is the correctness oracle for the
 * HIP implementation under nemaktis_amd/ and the CPU baseline timed by
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker / the baseline, never as the
 * thing measured or shipped.
 *
 * The restatement works on the FULL Nx x Ny grid with the reference's
 * duplicate-point semantics (column Nx-1 / row Ny-1 mirror column / row 0),
 * stores each sparse sub-matrix as explicit 3- or 4-point stencils, and
 * follows the reference's operation order.  Citations are relative to the
 * reference repository root.
 *
 * How each part is pinned (tests/test_oracle_pins.py): permittivity and
 * sqrt(K) by closed form and S*S = K; the phase operator by a 2x2 matrix
 * exponential; cyclic / block line solves by dense solves; the R
 * coefficients by second-order convergence to the continuum operator; the ADI
 * step by its exact Fourier symbol in homogeneous anisotropic media; the whole
 * march by x,y-invariant media (products of 2x2 matrices) and the uniform
 * wave plate (retardation k0 (ne - no) L, Beer-Lambert damping of the
 * extraordinary wave by ne_imag); Fresnel by the
 * textbook isotropic transmission and, for an in-plane uniaxial director, the
 * eigen-polarisation transmissions; the shift by translation and linear
 * precision; the FFT by numpy; the screen filter by the NA low-pass, a single
 * layer at normal incidence and the angular-spectrum propagator; Q-tensor
 * input by the Q of a director reproducing director mode; the focus
 * correction's exit index at 0.6 um by a layer index-matched to a dispersive
 * exit medium at another wavelength (the angular-spectrum propagator).  The assembled
 * march as a whole (tests/test_physics_uniaxial.py): a Gaussian beam through
 * a uniform tilted uniaxial crystal converges to the exact plane-wave
 * eigenmode propagation (the reference's own accuracy benchmark), walk-off
 * included.
 * The seam rule (+x/+y neighbours of the last march points read 1/eps_zz of
 * the duplicate column/row) is pinned in its structure: the duplicate plane
 * enters the march through its eps_zz and through nothing else.
 * Parity unpinned (no reference vectors and no closed form): the bit-level
 * composition of the R assembly in inhomogeneous media with walk-off, the
 * reference's Eigen summation order and FFTW rounding, the outcome of the
 * reference's racy W loop (the full sum is computed here) and the seam
 * coefficient's form, which is followed from reading the reference only.
 */
#ifndef NEMAKTIS_BPM_ORACLE_H
#define NEMAKTIS_BPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and numerical set-up of one oracle run.  Index values are already
 * evaluated per wavelength (the expression strings of the settings JSON are
 * host logic, tested on their own). */
typedef struct {
    int32_t Nx, Ny, Nz;
    int32_t field_dim;          /* 3 = director, 6 = Q-tensor (xx,yy,zz,xy,xz,yz) */
    double dx, dy, dz;          /* mesh spacings (um) */
    int32_t n_wl, n_q;
    const double *wl;           /* n_wl wavelengths (um) */
    const double *q;            /* 2*n_q incidence wavevectors (qx,qy) */
    const double *no, *ne, *ne_imag, *nhost, *nin, *nout; /* n_wl each */
    double nout_ref;            /* nout evaluated at 0.6 um (focus correction) */
    int32_t beam;               /* 0 UniformBeam, 1 GaussianBeam(waist), 2 None */
    double waist;
    int32_t n_woodbury;
    int32_t n_iso;
    const double *iso_thick, *iso_index; /* n_iso each */
    double z_foc, NA;
    const double *mask;         /* nullable, Nx*Ny*Nz raw values, >= 0 means LC */
    const double *weights;      /* nullable, 2*Nx*Ny*(Nz-1) interface weights */
    int32_t nthreads;           /* <= 0: all processors */
    int32_t apply_filter;       /* 1: run the screen filter (FFT) at the end */
} OracleParams;

/* Full propagation.  E is (n_wl, n_q, 4, Ny, Nx) complex128, interleaved
 * (re, im).  With beam == 2 it holds the caller's input fields on entry.
 * Returns 0 on success, a negative code on invalid input. */
int oracle_propagate(const OracleParams *p, const double *lc, double *E);
/* the same, also returning every transverse plane of the march (bulk output):
 * bulk[w][q][pol][plane][comp][m] complex128, planes 0..Nz-1 */
int oracle_propagate_bulk(const OracleParams *p, const double *lc, double *E, double *bulk);

/* ---- operator-level access for unit tests (one wavelength at a time) ---- */
void *oracle_create(const OracleParams *p, const double *lc);
void oracle_destroy(void *h);
/* computes nref for wavelength iwl and resets the slab counter */
int oracle_set_wavelength(void *h, int iwl);
double oracle_nref(void *h);
/* permittivity plane (12 values per point: eps xx,yy,zz,xy,xz,yz; K xx,yy,xy;
 * sqrt(K) xx,yy,xy) of plane iz, written to out[12*Nx*Ny] */
void oracle_eps_plane(void *h, int iz, double *out);
/* builds D1, D, R and the phase operator for slab iz (planes iz, iz+1) */
void oracle_update(void *h, int iz);
/* R coefficient of sub-matrix `type` (0 DX, 1 DY, 2 DXDY), block (0..3),
 * stencil slot `shift` (DX/DY: 0 minus, 1 centre, 2 plus; DXDY: 0 mxmy,
 * 1 pxmy, 2 mxpy, 3 pxpy), for every full-grid row point -> out[2*Nx*Ny] */
void oracle_get_R(void *h, int type, int block, int shift, double *out);
/* dst[dcomp] (+)= A_block * src[scomp] on one sub-matrix, with the
 * duplicate-point copies; fields are 2*Nx*Ny complex (interleaved) */
void oracle_block_vmult(void *h, int type, int block, double *dst, int dcomp,
                        const double *src, int scomp, int add);
/* dst[dcomp] = (I - mu A_block)^-1 src[scomp] along x (dir 0) or y (dir 1) */
void oracle_block_solve(void *h, int dir, int block, double *dst, int dcomp,
                        const double *src, int scomp);
/* single cyclic line solve (a_j x_{j-1} + b_j x_j + c_j x_{j+1} = d_j, indices
 * mod M), Sherman-Morrison / Thomas as the reference's line solver */
void oracle_cyclic_line_solve(int M, const double *a, const double *b,
                              const double *c, const double *d, double *x);
int oracle_apply_primary(void *h, double *field);     /* ADI step U */
void oracle_apply_phase(void *h, double *field);      /* P (half step) */
void oracle_apply_shift(void *h, int iq, double *field); /* Sh (half step) */
void oracle_phase_coefs(void *h, double *out);        /* op_xx, op_yy, op_xy */
/* screen filter of one wavelength on the 4*n_q planes of E_wl (in place) */
void oracle_filter_wavelength(void *h, int iwl, double *E_wl);
/* 2-D DFT of howmany planes (Ny x Nx), sign -1 forward / +1 backward, unnormalised */
void oracle_fft2(int Nx, int Ny, int howmany, int sign, double *data);

#ifdef __cplusplus
}
#endif

#endif
