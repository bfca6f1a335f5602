// Host-side launchers of the HIP kernels of the bpm-solver march.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace nmk {

// sum of n_eff over every voxel of each slot's wavelength -> sp[slot].nref
void launch_nref(const double *lc, const uint8_t *mask, const Geo &g, SlotParams *sp, int nslots, double *partials,
                 int nblocks, hipStream_t st);

// slab-averaged material data of slab iz for every slot (FP64, full grid)
void launch_slab_material(const double *lc, const uint8_t *mask, const double *weights, const Geo &g,
                          const SlotParams *sp, int nslots, int iz, double *sm, hipStream_t st);

// mu*R stencil coefficients and the phase operator on the march grid
template <class T>
void launch_coefs(const double *sm, const Geo &g, const SlotParams *sp, int nslots, C2<T> *cf, hipStream_t st);

// input beam + input Fresnel (beam != None) or packing of caller fields
template <class T>
void launch_init_fields(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                        const PropParams *pp, int nprops, int gaussian, double waist, double dx, double dy,
                        C2<T> *F, hipStream_t st);
template <class T>
void launch_pack_fields(const c64 *E, const Geo &g, const PropParams *pp, int nprops, int nq, C2<T> *F,
                        hipStream_t st);

template <class T>
void launch_shift(const C2<T> *in, C2<T> *out, const Geo &g, const PropParams *pp, int nprops, hipStream_t st);
template <class T>
void launch_phase(C2<T> *f, const C2<T> *cf, const Geo &g, const PropParams *pp, int nprops, hipStream_t st);
// out = in + mu*A_y in (dir 1) or in + mu*A_x in (dir 0)
template <class T>
void launch_explicit(int dir, const C2<T> *in, C2<T> *out, const C2<T> *cf, const Geo &g, const PropParams *pp,
                     int nprops, hipStream_t st);
// cyclic block-tridiagonal line solves along x (dir 0) or y (dir 1):
// neumann == 0: dst = M src, acc = dst;  neumann == 1: dst = M (mu O src), acc += dst
template <class T>
void launch_line_solve(int dir, int neumann, const C2<T> *src, C2<T> *dst, C2<T> *acc, const C2<T> *cf,
                       const Geo &g, const PropParams *pp, int nprops, hipStream_t st);
// interior march state -> full-grid complex128 fields with the global phase
// and the output Fresnel interface
template <class T>
void launch_output(const C2<T> *F, const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                   const PropParams *pp, int nprops, int nq, c64 *E, hipStream_t st);

// screen filter: batched 2-D FFT of howmany planes (Ny x Nx) and the per-k
// 2x2 transfer filter
struct FftPlan {
    int n, nst;
    int radix[24];
};
FftPlan make_fft_plan(int n);
bool fft_supported(int n);
void launch_fft2(c64 *data, int Nx, int Ny, int planes, int sign, const c64 *tw_x, const c64 *tw_y,
                 const FftPlan &px, const FftPlan &py, hipStream_t st);
void launch_filter(c64 *E, const Geo &g, const SlotParams *sp, int nslots, const double *qxy, int nq,
                   const double *iso_thick, const double *iso_index, int n_iso, double dx, double dy,
                   double z_foc, double NA, hipStream_t st);

}  // namespace nmk
