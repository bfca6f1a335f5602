// Host-side launchers of the HIP kernels of the bpm-solver march.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "expr_program.h"

namespace nmk {

// LC flags of the voxels from a float64 mask on the device (mask >= 0 is LC,
// UtilitiesLib/VectorField.cpp:61-62)
void launch_mask_flags(const double *mask, uint8_t *flags, int64_t n, hipStream_t st);
// interface weights of a compiled mask formula: w (2*Nx*Ny*(Nz-1) values),
// *mismatch |= 1 when the formula's crossings disagree with the LC flags
void launch_interface_weights(const ExprProgram &prog, const uint8_t *mask, int Nx, int Ny, int Nz, double dx,
                              double dy, double dz, double *w, int *mismatch, hipStream_t st);

// sum of n_eff over every voxel of each slot's wavelength -> sp[slot].nref
void launch_nref(const double *lc, const uint8_t *mask, const Geo &g, SlotParams *sp, int nslots, double *partials,
                 int nblocks, hipStream_t st);

// blocks of one slot's coefficient launch per slab (march-grid tiles)
int64_t slab_coef_tiles(int Mx, int My);

// slab material of slabs iz0 .. iz0+nz-1 (FP64) and from it the mu*R stencil
// coefficients and the phase operator on the march grid, for every slot, in
// one launch; slab iz writes buffer iz % ring of cf (cfstride elements per
// buffer) and of xt (the x-line tridiagonal rows in the chunk-interleaved
// layout of lpx, xtstride elements per buffer; xt may be null)
template <class T>
void launch_slab_coefs(const double *lc, const uint8_t *mask, const double *weights, const Geo &g,
                       const SlotParams *sp, int nslots, int iz0, int nz, C2<T> *cf, int64_t cfstride, C2<T> *xt,
                       int64_t xtstride, int ring, const LinePlan &lpx, hipStream_t st);

// input beam + input Fresnel (beam != None) or packing of caller fields
template <class T>
void launch_init_fields(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                        const PropParams *pp, int nprops, int gaussian, double waist, double dx, double dy,
                        C2<T> *F, hipStream_t st);
template <class T>
void launch_pack_fields(const c64 *E, const Geo &g, const PropParams *pp, int nprops, int nq, C2<T> *F,
                        hipStream_t st);

// slab boundary out = Pn Sh^nshift Pp in (cf_prev / cf_next may be null);
// unshifted: every propagation's shift filter is the identity (pointwise kernel)
template <class T>
void launch_boundary(const C2<T> *in, C2<T> *out, const C2<T> *cf_prev, const C2<T> *cf_next, const Geo &g,
                     const PropParams *pp, int nprops, int nshift, bool unshifted, hipStream_t st);
// out = in + mu*A_y in (dir 1) or in + mu*A_x in (dir 0)
template <class T>
void launch_explicit(int dir, const C2<T> *in, C2<T> *out, const C2<T> *cf, const Geo &g, const PropParams *pp,
                     int nprops, hipStream_t st);
// order: bit 0 puts the two components of an x row block next to each other
// in the XCD work ranges (else neighbouring row blocks), bit 1 the same for
// y strips, bit 2 puts the polarisation pairs innermost, bit 3 (before bit 2)
// puts the pairs of one group innermost and the groups outermost, the group
// size (pairs per wavelength slot) in bits 4 and up.
// lean: one gather point per round and no row prefetch (fewer registers,
// more waves per SIMD for large batches).  both (x lines): one block holds
// both field components of its rows (equal work per block).  pair (y lines): both polarisations
// in one block, gathered together (2*P*W threads).
// cyclic block-tridiagonal line solves of the implicit half-steps, one
// thread per chunk carrying both polarisations (nprops must be even, the
// propagations ordered pol + 2 * (wavevector + n_wavevectors * slot)):
//   x lines, neumann == 0: dst = M_x (I + mu A_y) src, acc = dst
//   x lines, neumann == 1: dst = M_x (mu O_x src),     acc += dst
//   y lines, neumann == 0: dst = M_y (I + mu A_x) src, acc = dst
//   y lines, neumann == 1: dst = M_y (mu O_y src),     acc += dst
LinePlan make_line_plan(int M, int chunks);
// LDS bytes one block may hold on the current device
size_t max_block_lds();
// raises the dynamic LDS limit of kernel f to lds bytes when above 64 KiB
void allow_kernel_lds(const void *f, size_t lds);
template <class T>
size_t xsolve_lds_bytes(const LinePlan &lp, int R, int nc);
template <class T>
size_t ysolve_lds_bytes(const LinePlan &lp, int W, int pair);
// Line solvers: y = M (RHS of src), then dst = y and acc = accin + y (neumann
// launches) or acc = y (first launch); a null dst or acc skips that store, and
// accin may alias src or acc
template <class T>
void launch_xsolve(int neumann, const C2<T> *src, C2<T> *dst, const C2<T> *accin, C2<T> *acc, const C2<T> *cf,
                   const C2<T> *xt, const Geo &g, const PropParams *pp, int nprops, const LinePlan &lp, int R,
                   int order, int lean, int both, hipStream_t st);
template <class T>
void launch_ysolve(int neumann, const C2<T> *src, C2<T> *dst, const C2<T> *accin, C2<T> *acc, const C2<T> *cf,
                   const Geo &g, const PropParams *pp, int nprops, const LinePlan &lp, int W, int order, int lean,
                   int pair, hipStream_t st);
// interior march state -> full-grid complex128 fields with the global phase
// and the output Fresnel interface
template <class T>
void launch_output(const C2<T> *F, const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                   const PropParams *pp, int nprops, int nq, c64 *E, hipStream_t st);

// output of a mesh of one plane: each full-grid point's entrance field (beam 0
// uniform, 1 Gaussian, 2 the caller's fields already in E) through the output
// interface, in place in E
void launch_output_entrance(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                            const PropParams *pp, int nprops, int beam, double waist, double dx, double dy, c64 *E,
                            hipStream_t st);
// plane 0 of the bulk output on the full grid: beam through the input
// interface (beam 0 uniform, 1 Gaussian) or the caller's fields Ein (2)
void launch_bulk_entrance(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                          const PropParams *pp, int nprops, int beam, double waist, double dx, double dy,
                          const c64 *Ein, int nq, c64 *bulk, hipStream_t st);
// one transverse plane of the bulk output: march state -> full grid with the
// phase of 2*plane shift half-steps, into
// bulk[(((iwl*nq + iq)*2 + pol)*Nz + plane)*2 + comp][m] (complex128)
template <class T>
void launch_bulk_plane(const C2<T> *F, const Geo &g, const SlotParams *sp, const PropParams *pp, int nprops,
                       int plane, int nq, c64 *bulk, hipStream_t st);

// screen filter: batched 2-D FFT of howmany planes (Ny x Nx) and the per-k
// 2x2 transfer filter
// radices of the stages, 4 bits each from the lowest (at most 16 stages)
// direct: the length has a prime factor above 13 or does not fit the LDS
// plans (more than 2048 points); its lines are transformed by a direct DFT
struct FftPlan {
    int n, nst;
    uint64_t radices;
    int direct;
};
FftPlan make_fft_plan(int n);
// true when the length factors into the radices of the LDS plans (2..13)
bool fft_supported(int n);
// unnormalised DFT of lines of length n, sign -1 forward, +1 backward, with
// the twiddle table tw[m] = exp(-2 pi i m / n): `lines` lines per plane, line
// i of plane p starting at p*plane_size + i*line_stride, elements elem_stride
// apart; src may equal dst
void launch_dft_lines(const c64 *src, c64 *dst, int n, int lines, int planes, int64_t plane_size,
                      int64_t line_stride, int64_t elem_stride, int sign, const c64 *tw, hipStream_t st);
void launch_fft2(c64 *data, int Nx, int Ny, int planes, int sign, const c64 *tw_x, const c64 *tw_y,
                 const FftPlan &px, const FftPlan &py, hipStream_t st);
// the two passes of launch_fft2: rows (src may equal dst) and columns in place
void launch_fft_rows(const c64 *src, c64 *dst, int Nx, int Ny, int planes, int sign, const c64 *tw_x,
                     const FftPlan &px, hipStream_t st);
void launch_fft_cols(c64 *data, int Nx, int Ny, int planes, int sign, const c64 *tw_y, const FftPlan &py,
                     hipStream_t st);
// adjacent columns per workgroup of the column pass (LDS 2*n*C complex128)
int fft_cols_per_block(int n);
void launch_filter(c64 *E, const Geo &g, const SlotParams *sp, int nslots, const double *qxy, int nq,
                   const double *iso_thick, const double *iso_index, int n_iso, double dx, double dy,
                   double z_foc, double NA, hipStream_t st);

}  // namespace nmk
