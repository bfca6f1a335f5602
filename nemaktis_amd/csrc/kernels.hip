// HIP kernels of the bpm-solver march for CDNA4 (gfx950).
//
// Data layout in HBM
//   director     lc[c + fd*(ix + Nx*(iy + Ny*iz))]   float64 (caller layout)
//   coefficients cf[slot*cf_slot_size(Mx, My) + cf_point(jx, jy, Mx) + c*64]  C2<T> (64-point
//                tiles: complex planes, then imaginary-only planes as T), march grid
//   march state  F[(prop*2 + comp)*NP + p]            C2<T>, march grid
//   output       E[(2*prop + comp)*Nfull + m]         complex128, full grid
// with m = ix + Nx*iy on the full grid and p = jx + Mx*jy on the march grid
// (the periodic interior without the duplicate column/row along periodic
// axes).  Every kernel walks points with consecutive threads on consecutive
// x so that loads and stores coalesce.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"
#include "fft_device.h"

#include <map>
#include <mutex>
#include <string>

namespace nmk {

#define NMK_LAUNCH_CHECK()                                                                       \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP launch failed: ") +     \
                                                       hipGetErrorString(e_));                   \
    } while (0)

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// permittivity of one voxel (PhysicsLib/PermittivityTensorField.cpp:18-93):
// e[0..5] eps xx,yy,zz,xy,xz,yz; e[6..8] K = transverse eps; e[9..11] sqrt(K),
// from the director (fd = 3) or Q-tensor (fd = 6) values of the voxel and its
// LC flag, fetched by load_voxel
// ---------------------------------------------------------------------------
struct VoxelData {
    double n[6];
    bool m;
};
__device__ __forceinline__ void load_voxel(const double *__restrict__ lc, const uint8_t *__restrict__ mask, int fd,
                                           int64_t v, VoxelData &d) {
    const double *n = lc + (int64_t)fd * v;
    d.m = mask ? (mask[v] != 0) : true;
    d.n[0] = n[0];
    d.n[1] = n[1];
    d.n[2] = n[2];
    if (fd == 6) {
        d.n[3] = n[3];
        d.n[4] = n[4];
        d.n[5] = n[5];
    }
}
__device__ __forceinline__ double eps_of(const VoxelData &d, int fd, const SlotParams &s, double *e) {
    const double *n = d.n;
    const bool m = d.m;
    if (fd == 3) {
        double n0 = n[0], n1 = n[1], n2 = n[2];
        e[0] = m ? s.eperp + s.ea * (n0 * n0) : s.ehost;
        e[1] = m ? s.eperp + s.ea * (n1 * n1) : s.ehost;
        e[2] = m ? s.eperp + s.ea * (n2 * n2) : s.ehost;
        e[3] = m ? s.ea * n0 * n1 : 0.0;
        e[4] = m ? s.ea * n0 * n2 : 0.0;
        e[5] = m ? s.ea * n1 * n2 : 0.0;
    } else {
        double ea_eff = 2 * s.ea / 3;
        double eiso = s.eperp + s.ea / 3.;
        e[0] = m ? eiso + ea_eff * n[0] : s.ehost;
        e[1] = m ? eiso + ea_eff * n[1] : s.ehost;
        e[2] = m ? eiso + ea_eff * n[2] : s.ehost;
        e[3] = m ? ea_eff * n[3] : 0.0;
        e[4] = m ? ea_eff * n[4] : 0.0;
        e[5] = m ? ea_eff * n[5] : 0.0;
    }
    e[6] = e[0] - (e[4] * e[4]) / e[2];
    e[7] = e[1] - (e[5] * e[5]) / e[2];
    e[8] = e[3] - e[4] * e[5] / e[2];
    double det = e[6] * e[7] - e[8] * e[8];
    double tr = 0.5 * (e[6] + e[7]);
    double neff;
    if (tr * tr - det < 1e-8) {
        neff = sqrt(tr);
    } else {
        double r = sqrt(tr * tr - det);
        neff = (sqrt(tr + r) + sqrt(tr - r)) / 2.;
    }
    e[9] = neff + (e[6] - e[7]) / (4 * neff);
    e[10] = neff + (e[7] - e[6]) / (4 * neff);
    e[11] = e[8] / (2 * neff);
    return neff;
}
__device__ __forceinline__ double eps_voxel(const double *__restrict__ lc, const uint8_t *__restrict__ mask, int fd,
                                            int64_t v, const SlotParams &s, double *e) {
    VoxelData d;
    load_voxel(lc, mask, fd, v, d);
    return eps_of(d, fd, s, e);
}

__global__ __launch_bounds__(256) void k_mask_flags(const double *__restrict__ mask, uint8_t *__restrict__ flags,
                                                    int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = mask[i] >= 0 ? 1 : 0;
}

void launch_mask_flags(const double *mask, uint8_t *flags, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_mask_flags, dim3(blocks), dim3(256), 0, st, mask, flags, n);
    NMK_LAUNCH_CHECK();
}

// Interface weights of a mask formula (UtilitiesLib/VectorField.cpp:64-113),
// one thread per (column, slab): the formula at the slab's two planes and, if
// it changes sign or vanishes between them, the crossing bracketed by
// bisection until the bracket is narrower than 1e-3 dz; w[2i], w[2i+1] of
// voxel i = ix + Nx (iy + Ny iz) become ((z - z_iz) / dz, 1 - that), other
// slabs keep 0.5.  A crossing between two planes of equal mask value sets
// *mismatch (the caller then drops every weight, as the reference does).
// The arithmetic is the host version's (settings.cpp), operation for
// operation; a plane's value is the same whichever slab evaluates it.
__global__ __launch_bounds__(256) void k_interface_weights(ExprProgram prog, const uint8_t *__restrict__ mask, int Nx,
                                                           int Ny, int Nz, double dx, double dy, double dz,
                                                           double *__restrict__ w, int *__restrict__ mismatch) {
#pragma clang fp contract(off)
    extern __shared__ double expr_stack[];  // [prog.depth][blockDim.x]
    const ExprStridedStack stk{expr_stack + threadIdx.x, (int)blockDim.x};
    const int64_t N = (int64_t)Nx * Ny, n = N * (Nz - 1);
    bool failed = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int iz = int(i / N);
        const int64_t iperp = i - (int64_t)iz * N;
        const int iy = int(iperp / Nx), ix = int(iperp - (int64_t)iy * Nx);
        const double x = (ix - 0.5 * (Nx - 1)) * dx, y = (iy - 0.5 * (Ny - 1)) * dy;
        auto f = [&](double z) { return expr_eval(prog, x, y, z, stk); };
        const double zi = (iz - 0.5 * (Nz - 1)) * dz;
        double zm = zi, zp = (iz + 1 - 0.5 * (Nz - 1)) * dz, z = zm;
        double fm = f(zm);
        const double fp = f(zp);
        double w0 = 0.5, w1 = 0.5;
        if (fm * fp <= 0) {
            const bool m0 = mask ? mask[i] != 0 : true, m1 = mask ? mask[i + N] != 0 : true;
            if (m0 == m1) failed = true;
            // each pass halves the bracket: about 10 passes
            for (int it = 0; it < 64 && fabs(zp - zm) >= 1e-3 * dz; it++) {
                z = zm;
                fm = f(z);
                z = 0.5 * (zp + zm);
                const double fz = f(z);
                if (fz == 0) break;
                if (fz * fm < 0) zp = z;
                else zm = z;
            }
            w0 = (z - zi) / dz;
            w1 = 1 - w0;
        }
        w[2 * i] = w0;
        w[2 * i + 1] = w1;
    }
    if (failed) atomicOr(mismatch, 1);
}

void launch_interface_weights(const ExprProgram &prog, const uint8_t *mask, int Nx, int Ny, int Nz, double dx,
                              double dy, double dz, double *w, int *mismatch, hipStream_t st) {
    const int64_t n = (int64_t)Nx * Ny * (Nz - 1);
    if (n <= 0) return;
    if (prog.depth < 1 || prog.depth > kExprMaxStack) throw std::runtime_error("interface weights: bad formula program");
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 18);
    const size_t lds = (size_t)prog.depth * 256 * sizeof(double);
    hipLaunchKernelGGL(k_interface_weights, dim3(blocks), dim3(256), lds, st, prog, mask, Nx, Ny, Nz, dx, dy, dz, w,
                       mismatch);
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// nref = mean n_eff over every voxel (PermittivityTensorField.cpp:94-98);
// fixed block ranges and fixed reduction trees make it deterministic
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nref_partial(const double *__restrict__ lc, const uint8_t *__restrict__ mask,
                                                      Geo g, const SlotParams *__restrict__ sp, double *partials,
                                                      int nblocks) {
    const int slot = blockIdx.y;
    const SlotParams s = sp[slot];
    const int64_t V = g.Nfull * g.Nz;
    const int64_t chunk = (V + nblocks - 1) / nblocks;
    const int64_t beg = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(V, beg + chunk);
    double acc = 0;
    double e[12];
    for (int64_t v = beg + threadIdx.x; v < end; v += blockDim.x) acc += eps_voxel(lc, mask, g.field_dim, v, s, e);
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[(int64_t)slot * nblocks + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_nref_final(SlotParams *sp, const double *__restrict__ partials, int nblocks,
                                                    double V) {
    const int slot = blockIdx.x;
    double acc = 0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) acc += partials[(int64_t)slot * nblocks + i];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sp[slot].nref = red[0] / V;
}

void launch_nref(const double *lc, const uint8_t *mask, const Geo &g, SlotParams *sp, int nslots, double *partials,
                 int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(k_nref_partial, dim3(nblocks, nslots), dim3(256), 0, st, lc, mask, g, sp, partials, nblocks);
    NMK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_nref_final, dim3(nslots), dim3(256), 0, st, sp, partials, nblocks,
                       (double)g.Nfull * (double)g.Nz);
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Slab material and mu*R coefficients of slab iz (planes iz, iz+1), one
// launch per slab for every slot.  A block takes a 32 x 8 tile of the march
// grid: its threads first evaluate, in FP64, the slab-averaged material of
// the tile and a one-point halo into LDS -- averaged 1/eps_zz, K,
// sqrt(K)^-1, summed eps_az/eps_zz -- and, at the tile's own points, the
// phase operator exp(i alpha d) (PhaseEvolutionOperator.cpp:34-63), stored
// beside the coefficients; then each thread assembles the coefficients of
// its point (ParaxialPrimaryEvolutionOperator.cpp:212-490) from the LDS
// tile:
//   D  = D0 + D1 K                      (D1 with the seam rule of :428-430)
//   R  = i W + (D S^-1 + S^-1 D) / 4,    W = (v_a d_b + d_a v_b) / 2
// for the three stencil families DX, DY, DXDY and their 2x2 blocks.  The
// halo holds the periodic neighbours of the march grid; 1/eps_zz of the
// duplicate column and row (full-grid index Nx-1, Ny-1), which the D1 seam
// rule reads at ix+1 = Mx or iy+1 = My, is evaluated separately.
// ---------------------------------------------------------------------------
// mu*R plane K and the following ones of one point, in the coefficient tile
// layout (complex planes at o, imaginary-only planes at oi); the plane
// indices are compile-time constants
template <class T, int K, int KEND>
__device__ __forceinline__ void store_mu_r(C2<T> *o, T *oi, const double (&re)[CF_NMU], const double (&im)[CF_NMU],
                                           double hz) {
    constexpr int idx = cf_index(K);
    if constexpr (cf_complex(K)) o[idx * kCfTile] = C2<T>{(T)(-hz * im[K]), (T)(hz * re[K])};
    else oi[idx * kCfTile] = (T)(hz * re[K]);
    if constexpr (K + 1 < KEND) store_mu_r<T, K + 1, KEND>(o, oi, re, im, hz);
}

constexpr int kMatTx = 32, kMatTy = 8, kMatHx = kMatTx + 2, kMatHy = kMatTy + 2, kMatHalo = kMatHx * kMatHy;

// slab averages of the fields the coefficients read, from the permittivity
// data a (plane iz) and b (plane iz+1) of one point
__device__ __forceinline__ void slab_fields(const double *a, const double *b, double *o) {
    o[SM_EINV] = (1. / a[2] + 1. / b[2]) / 2;
    o[SM_KXX] = (a[6] + b[6]) / 2;
    o[SM_KYY] = (a[7] + b[7]) / 2;
    o[SM_KXY] = (a[8] + b[8]) / 2;
    const double sxx = (a[9] + b[9]) / 2, syy = (a[10] + b[10]) / 2, sxy = (a[11] + b[11]) / 2;
    const double det = sxx * syy - sxy * sxy;
    o[SM_IXX] = syy / det;
    o[SM_IYY] = sxx / det;
    o[SM_IXY] = -sxy / det;
    o[SM_VX] = a[4] / a[2] + b[4] / b[2];
    o[SM_VY] = a[5] / a[2] + b[5] / b[2];
}

// phase operator exp(i alpha d) of one point, d = dZ/2 (w S(iz) + wpz S(iz+1) - no I)
__device__ __forceinline__ void slab_phase(const double *a, const double *b, double w, double wpz,
                                           const SlotParams &s, c64 &pxx, c64 &pyy, c64 &pxy) {
    const double dZ = s.dZ;
    double dxx = 0.5 * dZ * (w * a[9] + wpz * b[9] - s.no);
    double dyy = 0.5 * dZ * (w * a[10] + wpz * b[10] - s.no);
    double dxy = 0.5 * dZ * (w * a[11] + wpz * b[11]);
    double diff = (dxx - dyy) / 2., tr = (dxx + dyy) / 2.;
    double dt = sqrt(dxy * dxy + diff * diff);
    // alpha = 1 + i*k: sin/cos of the complex argument alpha*dt = (dt, k*dt)
    const double kk = s.alpha_im;
    double su, cu, st, ct;
    sincos(dt, &su, &cu);
    sincos(tr, &st, &ct);
    // the hyperbolic factors and the damping only in an absorbing medium (k is uniform per slot)
    double chv = 1.0, shv = 0.0, amp = 1.0;
    if (kk != 0.0) {
        const double vv = kk * dt;
        if (vv != 0.0) {
            chv = cosh(vv);
            shv = sinh(vv);
        }
        amp = exp(-kk * tr);
    }
    c64 sinc = fabs(dt) < 1e-8 ? mk(1.0, 0.0) : mk(su * chv / dt, cu * shv / dt);
    c64 cs = mk(cu * chv, -su * shv);
    // exp(i*alpha*tr) = exp(-k*tr) * (cos tr, sin tr)
    c64 ph = mk(amp * ct, amp * st);
    c64 ids = mk(-diff * sinc.y, diff * sinc.x);   // i*diff*sinc
    c64 ixs = mk(-dxy * sinc.y, dxy * sinc.x);     // i*dxy*sinc
    pxx = ph * (cs + ids);
    pyy = ph * (cs - ids);
    pxy = ph * ixs;
}

__device__ __forceinline__ int wrap_index(int v, int M) {
    v %= M;
    return v < 0 ? v + M : v;
}

// BATCH: one launch computes slabs iz0 .. iz0 + gridDim.y / nslots - 1
// (blockIdx.y = slab offset * nslots + slot), slab iz writing buffer iz % ring
// of the ring cf (cfstride elements per buffer) and of the x-line rows xt
// (xtstride); otherwise slab iz0 into cf and xt as given (blockIdx.y = slot).
template <class T, bool BATCH>
__global__ __launch_bounds__(256) void k_slab_coefs(const double *__restrict__ lc, const uint8_t *__restrict__ mask,
                                                    const double *__restrict__ weights, Geo g,
                                                    const SlotParams *__restrict__ sp, int iz0, int nslots,
                                                    C2<T> *__restrict__ cf, int64_t cfstride,
                                                    C2<T> *__restrict__ xt, int64_t xtstride, int ring,
                                                    LinePlan lpx) {
    __shared__ double mat[SM_COUNT * kMatHalo];
    __shared__ double seam_x[kMatTy], seam_y[kMatTx];
    const int ntx = (g.Mx + kMatTx - 1) / kMatTx, nty = (g.My + kMatTy - 1) / kMatTy;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= ntx * nty) return;
    const int ix0 = (tile % ntx) * kMatTx, iy0 = (tile / ntx) * kMatTy;
    const int slot = BATCH ? (int)blockIdx.y % nslots : (int)blockIdx.y;
    const int iz = BATCH ? iz0 + (int)blockIdx.y / nslots : iz0;
    const SlotParams s = sp[slot];
    const int64_t N = g.Nfull;
    if (BATCH) {
        cf += (iz % ring) * cfstride;
        if (xt) xt += (iz % ring) * xtstride;
    }
    C2<T> *cfs = cf + (size_t)slot * cf_slot_size(g.Mx, g.My);

    // material of the tile and its halo (periodic neighbours on the march
    // grid): a thread takes halo points t and t + 256 and fetches the voxels
    // of both before it evaluates either
    static_assert(kMatHalo <= 2 * 256, "two halo points per thread of a 256-thread block");
    int64_t hm[2];
    VoxelData va[2], vb[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int h = (int)threadIdx.x + 256 * r;
        if (h < kMatHalo) {
            const int lx = h % kMatHx, ly = h / kMatHx;
            const int jx = wrap_index(ix0 - 1 + lx, g.Mx), jy = wrap_index(iy0 - 1 + ly, g.My);
            hm[r] = jx + (int64_t)g.Nx * jy;
            load_voxel(lc, mask, g.field_dim, hm[r] + N * iz, va[r]);
            load_voxel(lc, mask, g.field_dim, hm[r] + N * (iz + 1), vb[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int h = (int)threadIdx.x + 256 * r;
        if (h >= kMatHalo) break;
        const int lx = h % kMatHx, ly = h / kMatHx;
        const int64_t m = hm[r];
        double a[12], b[12], f[SM_COUNT];
        eps_of(va[r], g.field_dim, s, a);
        eps_of(vb[r], g.field_dim, s, b);
        slab_fields(a, b, f);
#pragma unroll
        for (int c = 0; c < SM_COUNT; c++) mat[c * kMatHalo + h] = f[c];
        const int ox = ix0 - 1 + lx, oy = iy0 - 1 + ly;
        if (lx >= 1 && lx <= kMatTx && ly >= 1 && ly <= kMatTy && ox < g.Mx && oy < g.My) {
            double w = 0.5, wpz = 0.5;
            if (weights) {
                const int64_t v = m + N * iz;
                w = weights[2 * v];
                wpz = weights[2 * v + 1];
            }
            c64 pxx, pyy, pxy;
            slab_phase(a, b, w, wpz, s, pxx, pyy, pxy);
            C2<T> *o = cfs + cf_point(ox, oy, g.Mx);
            o[(CF_P + 0) * kCfTile] = C2<T>{(T)pxx.x, (T)pxx.y};
            o[(CF_P + 1) * kCfTile] = C2<T>{(T)pyy.x, (T)pyy.y};
            o[(CF_P + 2) * kCfTile] = C2<T>{(T)pxy.x, (T)pxy.y};
        }
    }
    // 1/eps_zz of the duplicate column (right of a tile ending at Mx-1) and row
    const int t = threadIdx.x;
    if (g.px && t < kMatTy && ix0 + kMatTx >= g.Mx && iy0 + t < g.My) {
        const int64_t m = g.Mx + (int64_t)g.Nx * (iy0 + t);
        double a[12], b[12];
        eps_voxel(lc, mask, g.field_dim, m + N * iz, s, a);
        eps_voxel(lc, mask, g.field_dim, m + N * (iz + 1), s, b);
        seam_x[t] = (1. / a[2] + 1. / b[2]) / 2;
    }
    if (g.py && t >= 64 && t < 64 + kMatTx && iy0 + kMatTy >= g.My && ix0 + t - 64 < g.Mx) {
        const int64_t m = (ix0 + t - 64) + (int64_t)g.Nx * g.My;
        double a[12], b[12];
        eps_voxel(lc, mask, g.field_dim, m + N * iz, s, a);
        eps_voxel(lc, mask, g.field_dim, m + N * (iz + 1), s, b);
        seam_y[t - 64] = (1. / a[2] + 1. / b[2]) / 2;
    }
    __syncthreads();

    const int ix = ix0 + t % kMatTx, iy = iy0 + t / kMatTx;
    if (ix >= g.Mx || iy >= g.My) return;
    const int64_t ip = ix + (int64_t)g.Mx * iy;
    // material at march point (jx, jy) -- a periodic neighbour of the tile --
    // or, for 1/eps_zz, at the duplicate column jx = Mx or row jy = My
    auto at = [&](int c, int jx, int jy) {
        if (c == SM_EINV) {
            if (jx == g.Mx) return seam_x[jy - iy0];
            if (jy == g.My) return seam_y[jx - ix0];
        }
        int hx = jx - ix0 + 1, hy = jy - iy0 + 1;
        if (hx < 0) hx += g.Mx;
        else if (hx > kMatTx + 1) hx -= g.Mx;
        if (hy < 0) hy += g.My;
        else if (hy > kMatTy + 1) hy -= g.My;
        return mat[c * kMatHalo + hx + kMatHx * hy];
    };
    const int xm = g.px ? (ix == 0 ? g.Mx - 1 : ix - 1) : ix;
    const int xp = g.px ? (ix == g.Mx - 1 ? 0 : ix + 1) : ix;
    const int ym = g.py ? (iy == 0 ? g.My - 1 : iy - 1) : iy;
    const int yp = g.py ? (iy == g.My - 1 ? 0 : iy + 1) : iy;

    double re[CF_NMU], im[CF_NMU];
#pragma unroll
    for (int k = 0; k < 40; k++) re[k] = im[k] = 0.0;

    const double dX = s.dX, dY = s.dY;
    const double ixx_r = at(SM_IXX, ix, iy), iyy_r = at(SM_IYY, ix, iy), ixy_r = at(SM_IXY, ix, iy);
    const double vx_r = at(SM_VX, ix, iy), vy_r = at(SM_VY, ix, iy);
    const double e_r = at(SM_EINV, ix, iy);

    // For each stencil slot: D of the four blocks, the walk-off part i*W, then
    // D.Sinv (Sinv at the column point) and Sinv.D (Sinv at the row point),
    // accumulated in the reference's order.
    // mu*R with mu = i*dZ/2 is stored family by family, so that only one
    // family's accumulators are live at a time
    const double hz = 0.5 * s.dZ;
    C2<T> *o = cfs + cf_point(ix, iy, g.Mx);
    T *oi = reinterpret_cast<T *>(o + CF_NCPLX * kCfTile) - cf_lane(ix, iy);
    if (g.px) {
        // ---- DX family: slots 0 minus, 1 centre, 2 plus ----
        const double e_mx = at(SM_EINV, xm, iy), e_px = at(SM_EINV, ix + 1, iy);  // ix+1: duplicate at the seam
        const double cm = 0.5 * (e_mx + e_r) / (dX * dX), cp = 0.5 * (e_px + e_r) / (dX * dX);
        const double d1[3] = {cm, -cm - cp, cp};
        const double d0[3] = {1. / (dX * dX), -2. / (dX * dX), 1. / (dX * dX)};
        const int qx[3] = {xm, ix, xp};
        for (int sl = 0; sl < 3; sl++) {
            const int q = qx[sl];
            double D[4] = {d1[sl] * at(SM_KXX, q, iy), d1[sl] * at(SM_KXY, q, iy), 0.0, d0[sl]};
            if (sl != 1) {
                const double d = (sl == 0 ? -1.0 : 1.0) * (0.5 / dX);
                const double vx_q = at(SM_VX, q, iy), vy_q = at(SM_VY, q, iy);
                im[CF_DX + 0 * 3 + sl] += d * (0.25 * vx_r);
                im[CF_DX + 2 * 3 + sl] += d * (0.25 * vy_r);
                im[CF_DX + 0 * 3 + sl] += d * (0.25 * vx_q);
                im[CF_DX + 1 * 3 + sl] += d * (0.25 * vy_q);
            }
            const double ixx = at(SM_IXX, q, iy), iyy = at(SM_IYY, q, iy), ixy = at(SM_IXY, q, iy);
            for (int rb = 0; rb < 2; rb++) {
                re[CF_DX + (rb * 2 + 0) * 3 + sl] += 0.25 * D[rb * 2 + 0] * ixx;
                re[CF_DX + (rb * 2 + 0) * 3 + sl] += 0.25 * D[rb * 2 + 1] * ixy;
                re[CF_DX + (rb * 2 + 1) * 3 + sl] += 0.25 * D[rb * 2 + 0] * ixy;
                re[CF_DX + (rb * 2 + 1) * 3 + sl] += 0.25 * D[rb * 2 + 1] * iyy;
            }
            for (int cb = 0; cb < 2; cb++) {
                re[CF_DX + (0 + cb) * 3 + sl] += 0.25 * ixx_r * D[0 + cb];
                re[CF_DX + (0 + cb) * 3 + sl] += 0.25 * ixy_r * D[2 + cb];
                re[CF_DX + (2 + cb) * 3 + sl] += 0.25 * ixy_r * D[0 + cb];
                re[CF_DX + (2 + cb) * 3 + sl] += 0.25 * iyy_r * D[2 + cb];
            }
        }
    }
    store_mu_r<T, CF_DX, CF_DY>(o, oi, re, im, hz);
    // (I - mu T) rows of the x-line systems (T = DX00, DX11) in the chunk-interleaved layout
    if (g.px && xt) {
        int pc, i;
        chunk_of(lpx, ix, pc, i);
        const int64_t SL = (int64_t)lpx.lmax * lpx.P;
        for (int c2 = 0; c2 < 2; c2++) {
            const int k0 = CF_DX + (c2 == 0 ? 0 : 3) * 3;
            const int64_t base = ((((int64_t)slot * 2 + c2) * 3) * g.My + iy) * SL + (int64_t)i * lpx.P + pc;
            const int64_t kst = (int64_t)g.My * SL;
            xt[base] = C2<T>{(T)(hz * im[k0 + 0]), (T)(-hz * re[k0 + 0])};
            xt[base + kst] = C2<T>{(T)(1.0 + hz * im[k0 + 1]), (T)(-hz * re[k0 + 1])};
            xt[base + 2 * kst] = C2<T>{(T)(hz * im[k0 + 2]), (T)(-hz * re[k0 + 2])};
        }
    }
    if (g.py) {
        // ---- DY family: slots 0 minus, 1 centre, 2 plus ----
        const double e_my = at(SM_EINV, ix, ym), e_py = at(SM_EINV, ix, iy + 1);  // iy+1: duplicate at the seam
        const double cm = 0.5 * (e_my + e_r) / (dY * dY), cp = 0.5 * (e_py + e_r) / (dY * dY);
        const double d1[3] = {cm, -cm - cp, cp};
        const double d0[3] = {1. / (dY * dY), -2. / (dY * dY), 1. / (dY * dY)};
        const int qy[3] = {ym, iy, yp};
        for (int sl = 0; sl < 3; sl++) {
            const int q = qy[sl];
            double D[4] = {d0[sl], 0.0, d1[sl] * at(SM_KXY, ix, q), d1[sl] * at(SM_KYY, ix, q)};
            if (sl != 1) {
                const double d = (sl == 0 ? -1.0 : 1.0) * (0.5 / dY);
                const double vx_q = at(SM_VX, ix, q), vy_q = at(SM_VY, ix, q);
                im[CF_DY + 1 * 3 + sl] += d * (0.25 * vx_r);
                im[CF_DY + 3 * 3 + sl] += d * (0.25 * vy_r);
                im[CF_DY + 2 * 3 + sl] += d * (0.25 * vx_q);
                im[CF_DY + 3 * 3 + sl] += d * (0.25 * vy_q);
            }
            const double ixx = at(SM_IXX, ix, q), iyy = at(SM_IYY, ix, q), ixy = at(SM_IXY, ix, q);
            for (int rb = 0; rb < 2; rb++) {
                re[CF_DY + (rb * 2 + 0) * 3 + sl] += 0.25 * D[rb * 2 + 0] * ixx;
                re[CF_DY + (rb * 2 + 0) * 3 + sl] += 0.25 * D[rb * 2 + 1] * ixy;
                re[CF_DY + (rb * 2 + 1) * 3 + sl] += 0.25 * D[rb * 2 + 0] * ixy;
                re[CF_DY + (rb * 2 + 1) * 3 + sl] += 0.25 * D[rb * 2 + 1] * iyy;
            }
            for (int cb = 0; cb < 2; cb++) {
                re[CF_DY + (0 + cb) * 3 + sl] += 0.25 * ixx_r * D[0 + cb];
                re[CF_DY + (0 + cb) * 3 + sl] += 0.25 * ixy_r * D[2 + cb];
                re[CF_DY + (2 + cb) * 3 + sl] += 0.25 * ixy_r * D[0 + cb];
                re[CF_DY + (2 + cb) * 3 + sl] += 0.25 * iyy_r * D[2 + cb];
            }
        }
    }
    store_mu_r<T, CF_DY, CF_XY>(o, oi, re, im, hz);
    if (g.px && g.py) {
        // ---- DXDY family: slots 0 mxmy, 1 pxmy, 2 mxpy, 3 pxpy ----
        const double dxy = dX * dY;
        const double e_px = at(SM_EINV, ix + 1, iy), e_mx = at(SM_EINV, xm, iy);
        const double e_py = at(SM_EINV, ix, iy + 1), e_my = at(SM_EINV, ix, ym);
        const double pp = 0.25 * e_px / dxy, mm = 0.25 * e_mx / dxy;
        const double qp = 0.25 * e_py / dxy, qm = 0.25 * e_my / dxy;
        const double i0 = 0.25 / dxy;
        const double m01[4] = {mm, -pp, -mm, pp};
        const double m10[4] = {qm, -qm, -qp, qp};
        const double dd[4] = {-i0, i0, i0, -i0};
        const int cx[4] = {xm, xp, xm, xp}, cy[4] = {ym, ym, yp, yp};
        for (int sl = 0; sl < 4; sl++) {
            const int qxx = cx[sl], qyy = cy[sl];
            const double kxx = at(SM_KXX, qxx, qyy), kyy = at(SM_KYY, qxx, qyy), kxy = at(SM_KXY, qxx, qyy);
            double D[4] = {m01[sl] * kxy, dd[sl] + m01[sl] * kyy, dd[sl] + m10[sl] * kxx, m10[sl] * kxy};
            const double ixx = at(SM_IXX, qxx, qyy), iyy = at(SM_IYY, qxx, qyy), ixy = at(SM_IXY, qxx, qyy);
            for (int rb = 0; rb < 2; rb++) {
                re[CF_XY + (rb * 2 + 0) * 4 + sl] += 0.25 * D[rb * 2 + 0] * ixx;
                re[CF_XY + (rb * 2 + 0) * 4 + sl] += 0.25 * D[rb * 2 + 1] * ixy;
                re[CF_XY + (rb * 2 + 1) * 4 + sl] += 0.25 * D[rb * 2 + 0] * ixy;
                re[CF_XY + (rb * 2 + 1) * 4 + sl] += 0.25 * D[rb * 2 + 1] * iyy;
            }
            for (int cb = 0; cb < 2; cb++) {
                re[CF_XY + (0 + cb) * 4 + sl] += 0.25 * ixx_r * D[0 + cb];
                re[CF_XY + (0 + cb) * 4 + sl] += 0.25 * ixy_r * D[2 + cb];
                re[CF_XY + (2 + cb) * 4 + sl] += 0.25 * ixy_r * D[0 + cb];
                re[CF_XY + (2 + cb) * 4 + sl] += 0.25 * iyy_r * D[2 + cb];
            }
        }
    }
    store_mu_r<T, CF_XY, CF_NMU>(o, oi, re, im, hz);
}

int64_t slab_coef_tiles(int Mx, int My) {
    return (int64_t)((Mx + kMatTx - 1) / kMatTx) * ((My + kMatTy - 1) / kMatTy);
}

template <class T>
void launch_slab_coefs(const double *lc, const uint8_t *mask, const double *weights, const Geo &g,
                       const SlotParams *sp, int nslots, int iz0, int nz, C2<T> *cf, int64_t cfstride, C2<T> *xt,
                       int64_t xtstride, int ring, const LinePlan &lpx, hipStream_t st) {
    if (nz < 1 || nz > ring || (int64_t)nz * nslots > 65535)
        throw std::runtime_error("slab coefficients: bad slab batch");
    const int64_t tiles = slab_coef_tiles(g.Mx, g.My);
    const dim3 grid((unsigned)((tiles + 7) / 8 * 8), nslots * nz);
    if (nz == 1) {
        C2<T> *x1 = xt ? xt + (iz0 % ring) * xtstride : nullptr;
        hipLaunchKernelGGL((k_slab_coefs<T, false>), grid, dim3(256), 0, st, lc, mask, weights, g, sp, iz0, nslots,
                           cf + (iz0 % ring) * cfstride, cfstride, x1, xtstride, ring, lpx);
    } else {
        hipLaunchKernelGGL((k_slab_coefs<T, true>), grid, dim3(256), 0, st, lc, mask, weights, g, sp, iz0, nslots,
                           cf, cfstride, xt, xtstride, ring, lpx);
    }
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// input beam (PhysicsLib/BeamProfile.cpp:17-39) + input Fresnel
// (BeamPropagationLib/FresnelOperator.cpp:16-38), written to the march grid
// ---------------------------------------------------------------------------
// beam of polarisation (p.ex, p.ey) at full-grid point (ix, iy) through the
// input interface
__device__ __forceinline__ void entrance_field(const double *__restrict__ lc, const uint8_t *__restrict__ mask,
                                               const Geo &g, const SlotParams &s, const PropParams &p, int ix,
                                               int iy, int gaussian, double waist, double dx, double dy,
                                               double &a, double &b) {
    double e[12];
    eps_voxel(lc, mask, g.field_dim, ix + (int64_t)g.Nx * iy, s, e);
    double gx = 1.0;
    if (gaussian) {
        double x = (ix - (g.Nx - 1) / 2.) * dx, y = (iy - (g.Ny - 1) / 2.) * dy;
        gx = exp(-(x * x + y * y) / (2 * waist * waist));
    }
    const double E0 = p.ex * gx, E1 = p.ey * gx;
    const double n = s.nin, nxx = e[9], nyy = e[10], nxy = e[11];
    const double det = (n + nxx) * (n + nyy) - nxy * nxy;
    a = 2 * n * (n + nyy) / det * E0 - 2 * n * nxy / det * E1;
    b = -2 * n * nxy / det * E0 + 2 * n * (n + nxx) / det * E1;
}

template <class T>
__global__ __launch_bounds__(256) void k_init_fields(const double *__restrict__ lc, const uint8_t *__restrict__ mask,
                                                     Geo g, const SlotParams *__restrict__ sp,
                                                     const PropParams *__restrict__ pp, int gaussian, double waist,
                                                     double dx, double dy, C2<T> *__restrict__ F) {
    const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= g.NP) return;
    const int prop = blockIdx.y;
    const PropParams p = pp[prop];
    const SlotParams s = sp[p.slot];
    const int ix = (int)(ip % g.Mx), iy = (int)(ip / g.Mx);
    double a, b;
    entrance_field(lc, mask, g, s, p, ix, iy, gaussian, waist, dx, dy, a, b);
    F[((int64_t)prop * 2 + 0) * g.NP + ip] = C2<T>{(T)a, (T)0};
    F[((int64_t)prop * 2 + 1) * g.NP + ip] = C2<T>{(T)b, (T)0};
}

// plane 0 of the bulk output on the full grid, duplicate column/row included
// (BPMIteration.cpp:88-120): the beam through the input interface at every
// point (beam 0 uniform, 1 Gaussian), or the caller's fields as given (2)
__global__ __launch_bounds__(256) void k_bulk_entrance(const double *__restrict__ lc,
                                                       const uint8_t *__restrict__ mask, Geo g,
                                                       const SlotParams *__restrict__ sp,
                                                       const PropParams *__restrict__ pp, int beam, double waist,
                                                       double dx, double dy, const c64 *__restrict__ Ein, int nq,
                                                       c64 *__restrict__ bulk) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.Nfull) return;
    const PropParams p = pp[blockIdx.y];
    const SlotParams s = sp[p.slot];
    c64 a, b;
    if (beam == 2) {
        const c64 *e = Ein + (((int64_t)s.iwl * nq + p.iq) * 4 + 2 * p.pol) * g.Nfull;
        a = e[m];
        b = e[g.Nfull + m];
    } else {
        double ar, br;
        entrance_field(lc, mask, g, s, p, (int)(m % g.Nx), (int)(m / g.Nx), beam == 1, waist, dx, dy, ar, br);
        a = mk(ar, 0.0);
        b = mk(br, 0.0);
    }
    c64 *o = bulk + (((int64_t)s.iwl * nq + p.iq) * 2 + p.pol) * g.Nz * 2 * g.Nfull;
    o[m] = a;
    o[g.Nfull + m] = b;
}

void launch_bulk_entrance(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                          const PropParams *pp, int nprops, int beam, double waist, double dx, double dy,
                          const c64 *Ein, int nq, c64 *bulk, hipStream_t st) {
    hipLaunchKernelGGL(k_bulk_entrance, dim3(blocks_for(g.Nfull, 256), nprops), dim3(256), 0, st, lc, mask, g, sp,
                       pp, beam, waist, dx, dy, Ein, nq, bulk);
    NMK_LAUNCH_CHECK();
}

template <class T>
void launch_init_fields(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                        const PropParams *pp, int nprops, int gaussian, double waist, double dx, double dy,
                        C2<T> *F, hipStream_t st) {
    hipLaunchKernelGGL(k_init_fields<T>, dim3(blocks_for(g.NP, 256), nprops), dim3(256), 0, st, lc, mask, g, sp, pp,
                       gaussian, waist, dx, dy, F);
    NMK_LAUNCH_CHECK();
}

template <class T>
__global__ __launch_bounds__(256) void k_pack_fields(const c64 *__restrict__ E, Geo g, C2<T> *__restrict__ F) {
    const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= g.NP) return;
    const int prop = blockIdx.y;
    const int ix = (int)(ip % g.Mx), iy = (int)(ip / g.Mx);
    const int64_t m = ix + (int64_t)g.Nx * iy;
    for (int c = 0; c < 2; c++) F[((int64_t)prop * 2 + c) * g.NP + ip] = cast2<T>(E[((int64_t)prop * 2 + c) * g.Nfull + m]);
}

template <class T>
void launch_pack_fields(const c64 *E, const Geo &g, const PropParams *pp, int nprops, int nq, C2<T> *F,
                        hipStream_t st) {
    (void)pp;
    (void)nq;
    hipLaunchKernelGGL(k_pack_fields<T>, dim3(blocks_for(g.NP, 256), nprops), dim3(256), 0, st, E, g, F);
    NMK_LAUNCH_CHECK();
}

__device__ __forceinline__ int wrapm(int i, int M) {
    int r = i % M;
    return r < 0 ? r + M : r;
}

// ---------------------------------------------------------------------------
// slab boundary: out = Pn . Sh^s . Pp (in).  Between slabs k and k+1 the
// reference applies P_k, Sh (end of slab k), Sh, P_{k+1} (start of slab k+1)
// (BPMIteration.cpp:135-144); the two Catmull-Rom shifts are composed into
// one 7-tap filter per axis (the x and y passes commute).  Pp / Pn are the
// phase operators of the previous / next slab (either may be absent), and
// the first and last boundaries carry a single shift.  One block computes a
// TW x TH output tile from an LDS copy of its (TW+6) x (TH+6) input window.
// ---------------------------------------------------------------------------
constexpr int kBTW = 64, kBTH = 8;

// One block: a TW x TH output tile of both polarisations of one (wavelength,
// wavevector) pair (they share the shift taps and the phase operators).  The
// (TW+6) x (TH+6) input window is loaded once with the previous phase operator
// applied, the x pass is done in place in LDS, then the y pass and the next
// phase operator; wrapped window indices are computed once per block.
template <class T>
__global__ __launch_bounds__(256) void k_boundary(const C2<T> *__restrict__ in, C2<T> *__restrict__ out,
                                                  const C2<T> *__restrict__ cf_prev, const C2<T> *__restrict__ cf_next,
                                                  Geo g, const PropParams *__restrict__ pp, int nshift) {
    using C = C2<T>;
    constexpr int WW = kBTW + 6, WH = kBTH + 6;
    __shared__ C tile[4][WH][kBTW + 7];
    __shared__ int colw[WW], roww[WH];
    const int prop0 = blockIdx.z * 2;
    const PropParams p = pp[prop0];
    const size_t NP = (size_t)g.NP;
    const int x0 = blockIdx.x * kBTW, y0 = blockIdx.y * kBTH;
    const double *wxs = nshift == 2 ? p.w2x : p.w1x, *wys = nshift == 2 ? p.w2y : p.w1y;
    const int ox = g.px ? (nshift == 2 ? p.o2x : p.o1x) : 0;
    const int oy = g.py ? (nshift == 2 ? p.o2y : p.o1y) : 0;
    if ((int)threadIdx.x < WW) {
        const int i = x0 + ox + (int)threadIdx.x;
        colw[threadIdx.x] = g.px ? wrapm(i, g.Mx) : min(max(i, 0), g.Mx - 1);
    } else if ((int)threadIdx.x < WW + WH) {
        const int i = y0 + oy + (int)threadIdx.x - WW;
        roww[threadIdx.x - WW] = g.py ? wrapm(i, g.My) : min(max(i, 0), g.My - 1);
    }
    __syncthreads();
    const C *f = in + (size_t)prop0 * 2 * NP;  // planes: pol 0 (x, y), pol 1 (x, y)
    const C *cp = cf_prev ? cf_prev + (size_t)p.slot * cf_slot_size(g.Mx, g.My) + CF_P * kCfTile : nullptr;
    // input window with the previous slab's phase operator applied: a
    // thread takes elements t, t + 256, ... and fetches all of them (fields
    // and phase operators) before it evaluates any
    constexpr int NWE = (WH * WW + 255) / 256;
    C wf[NWE][4], wp[NWE][3];
#pragma unroll
    for (int k = 0; k < NWE; k++) {
        const int e = threadIdx.x + k * 256;
        if (e < WH * WW) {
            const int r = e / WW, c = e - r * WW;
            const unsigned q = colw[c] + (unsigned)g.Mx * roww[r];
#pragma unroll
            for (int pl = 0; pl < 4; pl++) wf[k][pl] = f[pl * NP + q];
            if (cp) {
                const C *c3 = cp + cf_point(colw[c], roww[r], g.Mx);
                wp[k][0] = c3[0];
                wp[k][1] = c3[kCfTile];
                wp[k][2] = c3[2 * kCfTile];
            }
        }
    }
    // the next slab's phase operators at the thread's output points t and
    // t + 256, fetched with the window
    const C *cn = cf_next ? cf_next + (size_t)p.slot * cf_slot_size(g.Mx, g.My) + CF_P * kCfTile : nullptr;
    static_assert(kBTH * kBTW == 2 * 256, "two output points per thread of a 256-thread block");
    C np3[2][3];
    if (cn) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int e = threadIdx.x + k * 256;
            const int r = e / kBTW, c = e - r * kBTW;
            const int x = x0 + c, y = y0 + r;
            if (x < g.Mx && y < g.My) {
                const C *c3 = cn + cf_point(x, y, g.Mx);
                np3[k][0] = c3[0];
                np3[k][1] = c3[kCfTile];
                np3[k][2] = c3[2 * kCfTile];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NWE; k++) {
        const int e = threadIdx.x + k * 256;
        if (e < WH * WW) {
            const int r = e / WW, c = e - r * WW;
            C a0 = wf[k][0], b0 = wf[k][1], a1 = wf[k][2], b1 = wf[k][3];
            if (cp) {
                const C pxx = wp[k][0], pyy = wp[k][1], pxy = wp[k][2];
                const C na0 = pxx * a0 + pxy * b0, nb0 = pxy * a0 + pyy * b0;
                const C na1 = pxx * a1 + pxy * b1, nb1 = pxy * a1 + pyy * b1;
                a0 = na0;
                b0 = nb0;
                a1 = na1;
                b1 = nb1;
            }
            tile[0][r][c] = a0;
            tile[1][r][c] = b0;
            tile[2][r][c] = a1;
            tile[3][r][c] = b1;
        }
    }
    __syncthreads();
    T wx[7], wy[7];
#pragma unroll
    for (int t = 0; t < 7; t++) {
        wx[t] = g.px ? (T)wxs[t] : (T)(t == 0 ? 1 : 0);
        wy[t] = g.py ? (T)wys[t] : (T)(t == 0 ? 1 : 0);
    }
    // x pass in place: every thread computes its elements into registers,
    // then all of them overwrite the window after a barrier
    constexpr int NXE = (WH * kBTW + 255) / 256;
    C xr[NXE][4];
#pragma unroll
    for (int k = 0; k < NXE; k++) {
        const int e = threadIdx.x + k * 256;
        if (e < WH * kBTW) {
            const int r = e / kBTW, c = e - r * kBTW;
#pragma unroll
            for (int pl = 0; pl < 4; pl++) {
                C a = wx[0] * tile[pl][r][c];
#pragma unroll
                for (int t = 1; t < 7; t++) a = a + wx[t] * tile[pl][r][c + t];
                xr[k][pl] = a;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NXE; k++) {
        const int e = threadIdx.x + k * 256;
        if (e < WH * kBTW) {
            const int r = e / kBTW, c = e - r * kBTW;
#pragma unroll
            for (int pl = 0; pl < 4; pl++) tile[pl][r][c] = xr[k][pl];
        }
    }
    __syncthreads();
    // y pass and the next slab's phase operator
    C *o = out + (size_t)prop0 * 2 * NP;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int e = threadIdx.x + k * 256;
        const int r = e / kBTW, c = e - r * kBTW;
        const int x = x0 + c, y = y0 + r;
        if (x >= g.Mx || y >= g.My) continue;
        C v[4];
#pragma unroll
        for (int pl = 0; pl < 4; pl++) {
            C a = wy[0] * tile[pl][r][c];
#pragma unroll
            for (int t = 1; t < 7; t++) a = a + wy[t] * tile[pl][r + t][c];
            v[pl] = a;
        }
        const unsigned q = x + (unsigned)g.Mx * y;
        if (cn) {
            const C pxx = np3[k][0], pyy = np3[k][1], pxy = np3[k][2];
#pragma unroll
            for (int pol = 0; pol < 2; pol++) {
                const C a = pxx * v[2 * pol] + pxy * v[2 * pol + 1], b = pxy * v[2 * pol] + pyy * v[2 * pol + 1];
                v[2 * pol] = a;
                v[2 * pol + 1] = b;
            }
        }
        o[q] = v[0];
        o[NP + q] = v[1];
        o[2 * NP + q] = v[2];
        o[3 * NP + q] = v[3];
    }
}

// The same boundary when no propagation of the launch is shifted (normal
// incidence: the Catmull-Rom filters are the identity), point by point
template <class T>
__global__ __launch_bounds__(256) void k_boundary_point(const C2<T> *__restrict__ in, C2<T> *__restrict__ out,
                                                        const C2<T> *__restrict__ cf_prev,
                                                        const C2<T> *__restrict__ cf_next, Geo g,
                                                        const PropParams *__restrict__ pp) {
    using C = C2<T>;
    const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (unsigned)g.NP) return;
    const int prop0 = blockIdx.y * 2;
    const int slot = pp[prop0].slot;
    const size_t NP = (size_t)g.NP;
    const C *f = in + (size_t)prop0 * 2 * NP;
    C v[4] = {f[q], f[NP + q], f[2 * NP + q], f[3 * NP + q]};
    const unsigned qy = q / (unsigned)g.Mx, qx = q - qy * (unsigned)g.Mx;
    auto apply = [&](const C *cfx) {
        const C *c3 = cfx + (size_t)slot * cf_slot_size(g.Mx, g.My) + CF_P * kCfTile + cf_point(qx, qy, g.Mx);
        const C pxx = c3[0], pyy = c3[kCfTile], pxy = c3[2 * kCfTile];
#pragma unroll
        for (int pol = 0; pol < 2; pol++) {
            const C a = pxx * v[2 * pol] + pxy * v[2 * pol + 1], b = pxy * v[2 * pol] + pyy * v[2 * pol + 1];
            v[2 * pol] = a;
            v[2 * pol + 1] = b;
        }
    };
    if (cf_prev) apply(cf_prev);
    if (cf_next) apply(cf_next);
    C *o = out + (size_t)prop0 * 2 * NP;
    o[q] = v[0];
    o[NP + q] = v[1];
    o[2 * NP + q] = v[2];
    o[3 * NP + q] = v[3];
}

template <class T>
void launch_boundary(const C2<T> *in, C2<T> *out, const C2<T> *cf_prev, const C2<T> *cf_next, const Geo &g,
                     const PropParams *pp, int nprops, int nshift, bool unshifted, hipStream_t st) {
    if (nprops % 2) throw std::runtime_error("slab boundary: propagations come in polarisation pairs");
    if (unshifted) {
        dim3 grid((unsigned)blocks_for(g.NP, 256), nprops / 2);
        hipLaunchKernelGGL(k_boundary_point<T>, grid, dim3(256), 0, st, in, out, cf_prev, cf_next, g, pp);
    } else {
        dim3 grid((g.Mx + kBTW - 1) / kBTW, (g.My + kBTH - 1) / kBTH, nprops / 2);
        hipLaunchKernelGGL(k_boundary<T>, grid, dim3(256), 0, st, in, out, cf_prev, cf_next, g, pp, nshift);
    }
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// transverse operator taps (ParaxialPrimaryEvolutionOperator.cpp:102-210).
// Every right-hand side of the z-march is a field value plus a sum of
// (mu*R coefficient) x (neighbouring field value) taps:
//   yplus  (I + mu A_y) X,  A_y = [DY00 DY01; DY10+XY10 DY11+XY11]
//   xplus  (I + mu A_x) S,  A_x = [DX00+XY00 DX01+XY01; DX10 DX11]
//   ox     mu O_x S,        O_x = [XY00  DX01+XY01; DX10  0]
//   oy     mu O_y S,        O_y = [0  DY01; DY10+XY10  XY11]
// (mu is folded into the stored coefficients).  A tap is (coefficient plane,
// field component, dx, dy); D? stencils are {minus, centre, plus} along their
// axis and XY stencils {(-,-), (+,-), (-,+), (+,+)}.
// ---------------------------------------------------------------------------
enum RhsKind : int { RHS_YPLUS = 0, RHS_XPLUS = 1, RHS_OX = 2, RHS_OY = 3 };

struct Tap {
    int k, fc, dx, dy;
};
struct TapList {
    int n, centre;  // centre: field component added with weight 1, or -1
    Tap t[14];
};

__host__ __device__ constexpr void push_d(TapList &L, int base, int blk, int fc, bool along_x) {
    for (int s = 0; s < 3; s++) {
        L.t[L.n] = along_x ? Tap{base + blk * 3 + s, fc, s - 1, 0} : Tap{base + blk * 3 + s, fc, 0, s - 1};
        L.n++;
    }
}
__host__ __device__ constexpr void push_xy(TapList &L, int blk, int fc) {
    constexpr int dx[4] = {-1, 1, -1, 1}, dy[4] = {-1, -1, 1, 1};
    for (int s = 0; s < 4; s++) {
        L.t[L.n] = Tap{CF_XY + blk * 4 + s, fc, dx[s], dy[s]};
        L.n++;
    }
}
__host__ __device__ constexpr TapList rhs_taps(int kind, int comp) {
    TapList L{0, -1, {}};
    if (kind == RHS_YPLUS) {
        L.centre = comp;
        if (comp == 0) {
            push_d(L, CF_DY, 0, 0, false);
            push_d(L, CF_DY, 1, 1, false);
        } else {
            push_d(L, CF_DY, 2, 0, false);
            push_xy(L, 2, 0);
            push_d(L, CF_DY, 3, 1, false);
            push_xy(L, 3, 1);
        }
    } else if (kind == RHS_XPLUS) {
        L.centre = comp;
        if (comp == 0) {
            push_d(L, CF_DX, 0, 0, true);
            push_xy(L, 0, 0);
            push_d(L, CF_DX, 1, 1, true);
            push_xy(L, 1, 1);
        } else {
            push_d(L, CF_DX, 2, 0, true);
            push_d(L, CF_DX, 3, 1, true);
        }
    } else if (kind == RHS_OX) {
        if (comp == 0) {
            push_xy(L, 0, 0);
            push_d(L, CF_DX, 1, 1, true);
            push_xy(L, 1, 1);
        } else {
            push_d(L, CF_DX, 2, 0, true);
        }
    } else {
        if (comp == 0) {
            push_d(L, CF_DY, 1, 1, false);
        } else {
            push_d(L, CF_DY, 2, 0, false);
            push_xy(L, 2, 0);
            push_xy(L, 3, 1);
        }
    }
    return L;
}

// periodic neighbour columns of a point and the element offsets Mx*row of
// its own row and the rows above and below (one multiply per point)
struct Nbr {
    unsigned xm, xp, rm, r0, rp;
};
__device__ __forceinline__ Nbr neighbours(const Geo &g, unsigned ix, unsigned iy) {
    Nbr n;
    const unsigned Mx = g.Mx;
    n.xm = g.px ? (ix == 0 ? Mx - 1 : ix - 1) : ix;
    n.xp = g.px ? (ix == Mx - 1 ? 0 : ix + 1) : ix;
    n.r0 = Mx * iy;
    n.rm = g.py ? (iy == 0 ? (unsigned)g.NP - Mx : n.r0 - Mx) : n.r0;
    n.rp = g.py ? (iy == (unsigned)g.My - 1 ? 0 : n.r0 + Mx) : n.r0;
    return n;
}

// acc + c f for the coefficient of plane K (imaginary-only planes: c = i c.y)
template <int K, class T>
__device__ __forceinline__ C2<T> tap_mad(C2<T> acc, C2<T> c, C2<T> f) {
    if constexpr (cf_complex(K)) return cmad(acc, c, f);
    else return C2<T>{acc.x - c.y * f.y, acc.y + c.y * f.x};
}

// Gathered operands of one right-hand side: every coefficient and field
// value of the point is loaded before any arithmetic, so that the loads of a
// point overlap in one memory round trip.  Indices are 32-bit element
// offsets from uniform plane pointers.
template <class T, int KIND, int COMP>
struct RhsGather {
    static constexpr int N = rhs_taps(KIND, COMP).n;
    static constexpr int CTR = rhs_taps(KIND, COMP).centre;
    C2<T> c[N], f[N], ctr;

    // tap I and the following ones; the tap table is evaluated at compile time
    template <int I>
    __device__ __forceinline__ void load_taps(const C2<T> *__restrict__ pc, unsigned lane,
                                              const C2<T> *__restrict__ f0, const C2<T> *__restrict__ f1,
                                              const Nbr &n, unsigned jx) {
        constexpr Tap tp = rhs_taps(KIND, COMP).t[I];
        const unsigned col = tp.dx < 0 ? n.xm : tp.dx > 0 ? n.xp : jx;
        const unsigned row = tp.dy < 0 ? n.rm : tp.dy > 0 ? n.rp : n.r0;
        c[I] = cf_load<tp.k>(pc, lane);
        f[I] = (tp.fc == 0 ? f0 : f1)[col + row];
        if constexpr (I + 1 < N) load_taps<I + 1>(pc, lane, f0, f1, n, jx);
    }
    template <int I>
    __device__ __forceinline__ C2<T> sum_taps(C2<T> a) const {
        a = tap_mad<rhs_taps(KIND, COMP).t[I].k>(a, c[I], f[I]);
        if constexpr (I + 1 < N) return sum_taps<I + 1>(a);
        else return a;
    }

    __device__ __forceinline__ void load(unsigned jx, unsigned jy, const C2<T> *__restrict__ cfs,
                                         const C2<T> *__restrict__ f0, const C2<T> *__restrict__ f1, const Geo &g,
                                         unsigned NP) {
        (void)NP;
        const Nbr n = neighbours(g, jx, jy);
        const unsigned ip = jx + n.r0;
        load_taps<0>(cfs + cf_point(jx, jy, g.Mx), cf_lane(jx, jy), f0, f1, n, jx);
        if constexpr (CTR >= 0) ctr = (CTR == 0 ? f0 : f1)[ip];
    }
    __device__ __forceinline__ C2<T> sum() const {
        C2<T> a = C2<T>{0, 0};
        if constexpr (CTR >= 0) a = ctr;
        return sum_taps<0>(a);
    }
};

// Both polarisations of one (wavelength, wavevector) read the same
// coefficients: one gather loads the coefficient taps once and the field
// taps of both.  a0/a1: components of polarisation 0, b0/b1: polarisation 1.
template <class T, int KIND, int COMP>
struct RhsGather2 {
    static constexpr int N = rhs_taps(KIND, COMP).n;
    static constexpr int CTR = rhs_taps(KIND, COMP).centre;
    C2<T> c[N], fa[N], fb[N], ca, cb;

    template <int I>
    __device__ __forceinline__ void load_taps(const C2<T> *__restrict__ pc, unsigned lane,
                                              const C2<T> *__restrict__ a0, const C2<T> *__restrict__ a1,
                                              const C2<T> *__restrict__ b0, const C2<T> *__restrict__ b1,
                                              const Nbr &n, unsigned jx) {
        constexpr Tap tp = rhs_taps(KIND, COMP).t[I];
        const unsigned col = tp.dx < 0 ? n.xm : tp.dx > 0 ? n.xp : jx;
        const unsigned q = col + (tp.dy < 0 ? n.rm : tp.dy > 0 ? n.rp : n.r0);
        c[I] = cf_load<tp.k>(pc, lane);
        fa[I] = (tp.fc == 0 ? a0 : a1)[q];
        fb[I] = (tp.fc == 0 ? b0 : b1)[q];
        if constexpr (I + 1 < N) load_taps<I + 1>(pc, lane, a0, a1, b0, b1, n, jx);
    }
    template <int I>
    __device__ __forceinline__ void sum_taps(C2<T> &x, C2<T> &y) const {
        constexpr int K = rhs_taps(KIND, COMP).t[I].k;
        x = tap_mad<K>(x, c[I], fa[I]);
        y = tap_mad<K>(y, c[I], fb[I]);
        if constexpr (I + 1 < N) sum_taps<I + 1>(x, y);
    }
    __device__ __forceinline__ void load(unsigned jx, unsigned jy, const C2<T> *__restrict__ cfs,
                                         const C2<T> *__restrict__ a0, const C2<T> *__restrict__ a1,
                                         const C2<T> *__restrict__ b0, const C2<T> *__restrict__ b1, const Geo &g) {
        const Nbr n = neighbours(g, jx, jy);
        const unsigned ip = jx + n.r0;
        load_taps<0>(cfs + cf_point(jx, jy, g.Mx), cf_lane(jx, jy), a0, a1, b0, b1, n, jx);
        if constexpr (CTR >= 0) {
            ca = (CTR == 0 ? a0 : a1)[ip];
            cb = (CTR == 0 ? b0 : b1)[ip];
        }
    }
    __device__ __forceinline__ void sum(C2<T> &ra, C2<T> &rb) const {
        C2<T> x{0, 0}, y{0, 0};
        if constexpr (CTR >= 0) {
            x = ca;
            y = cb;
        }
        sum_taps<0>(x, y);
        ra = x;
        rb = y;
    }
};

// field planes of one polarisation pair: prop0 = 2 * pair
template <class T>
struct PairFields {
    const C2<T> *a0, *a1, *b0, *b1;
    __device__ __forceinline__ PairFields(const C2<T> *src, int prop0, size_t NP) {
        a0 = src + (size_t)prop0 * 2 * NP;
        a1 = a0 + NP;
        b0 = a1 + NP;
        b1 = b0 + NP;
    }
};

// points per thread per gather round: four in FP32 (two beside the registers
// of 16-point chunks), one in FP64; stencils wider than 8 taps gather half of
// them at a time
template <class T, int LMAX>
constexpr int rhs_points() {
    return sizeof(T) == 4 ? (LMAX <= 8 ? 4 : 2) : 1;
}

// right-hand sides of component `comp` (uniform per block) for both
// polarisations at NPR points (v[k] false: skipped); every load of a group of
// points is issued before any sum
template <class T, int KIND, int COMP, int NPR>
__device__ __forceinline__ void rhs_pair_comp(const unsigned (&xs)[NPR], const unsigned (&ys)[NPR],
                                              const bool (&v)[NPR], const C2<T> *__restrict__ cfs,
                                              const PairFields<T> &f, const Geo &g, C2<T> (&r)[NPR][2]) {
    constexpr int G = RhsGather2<T, KIND, COMP>::N > 8 && NPR > 1 ? NPR / 2 : NPR;
#pragma unroll
    for (int k0 = 0; k0 < NPR; k0 += G) {
        RhsGather2<T, KIND, COMP> gs[G];
#pragma unroll
        for (int k = 0; k < G; k++)
            if (v[k0 + k]) gs[k].load(xs[k0 + k], ys[k0 + k], cfs, f.a0, f.a1, f.b0, f.b1, g);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < G; k++)
            if (v[k0 + k]) gs[k].sum(r[k0 + k][0], r[k0 + k][1]);
    }
}
template <class T, int KIND, int NPR>
__device__ __forceinline__ void rhs_pair_n(int comp, const unsigned (&xs)[NPR], const unsigned (&ys)[NPR],
                                           const bool (&v)[NPR], const C2<T> *__restrict__ cfs,
                                           const PairFields<T> &f, const Geo &g, C2<T> (&r)[NPR][2]) {
    if (comp == 0) rhs_pair_comp<T, KIND, 0, NPR>(xs, ys, v, cfs, f, g, r);
    else rhs_pair_comp<T, KIND, 1, NPR>(xs, ys, v, cfs, f, g, r);
}

// explicit half-steps, used along an axis too short for line solves
// (ParaxialPrimaryEvolutionOperator.cpp:153-160, 181-188):
//   dir 1: out = (I + mu A_y) in,   dir 0: out = (I + mu A_x) in
template <class T, int DIR>
__global__ __launch_bounds__(256) void k_explicit(const C2<T> *__restrict__ in, C2<T> *__restrict__ out,
                                                  const C2<T> *__restrict__ cf, Geo g,
                                                  const PropParams *__restrict__ pp) {
    const unsigned ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= (unsigned)g.NP) return;
    const int prop = blockIdx.y;
    const int slot = pp[prop].slot;
    const unsigned NP = (unsigned)g.NP, ix = ip % (unsigned)g.Mx, iy = ip / (unsigned)g.Mx;
    const C2<T> *c = cf + (size_t)slot * cf_slot_size(g.Mx, g.My);
    const C2<T> *f0 = in + ((size_t)prop * 2) * NP, *f1 = f0 + NP;
    constexpr int KIND = DIR == 1 ? RHS_YPLUS : RHS_XPLUS;
    RhsGather<T, KIND, 0> a;
    RhsGather<T, KIND, 1> b;
    a.load(ix, iy, c, f0, f1, g, NP);
    b.load(ix, iy, c, f0, f1, g, NP);
    __builtin_amdgcn_sched_barrier(0);
    C2<T> *o0 = out + ((size_t)prop * 2) * NP;
    o0[ip] = a.sum();
    o0[NP + ip] = b.sum();
}

template <class T>
void launch_explicit(int dir, const C2<T> *in, C2<T> *out, const C2<T> *cf, const Geo &g, const PropParams *pp,
                     int nprops, hipStream_t st) {
    dim3 grid(blocks_for(g.NP, 256), nprops);
    if (dir == 0) hipLaunchKernelGGL((k_explicit<T, 0>), grid, dim3(256), 0, st, in, out, cf, g, pp);
    else hipLaunchKernelGGL((k_explicit<T, 1>), grid, dim3(256), 0, st, in, out, cf, g, pp);
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// cyclic line solves (I - mu T) x = d, T = DX00/DX11 along x or DY00/DY11
// along y (line solver of BlockSparseMatrix.cpp:194-258, 385-450).
//
// A block holds a set of lines for both polarisations of one (wavelength,
// wavevector): their right-hand sides are gathered together (the two
// polarisations share every coefficient load) and staged in LDS, then each
// line is cut into P chunks (P a power of two) and one thread owns one chunk
// of one polarisation: a Thomas elimination of the chunk interior for the
// right-hand side and for the two coupling responses, then the P
// chunk-boundary unknowns form a cyclic tridiagonal system solved by
// parallel cyclic reduction in LDS, and the thread rebuilds its chunk.
// ---------------------------------------------------------------------------
// Work item `it` of a solver launch as (pair, rest) when the pairs come in
// groups of G (the wavevector pairs of one wavelength) and each group sweeps
// the nrest row blocks / strips with its G pairs innermost, so that one
// wavelength's coefficient rows serve its pairs while they are in L2
__device__ __forceinline__ void split_grouped(int it, int nrest, int npairs, int G, int &pair, int &rest) {
    G = max(1, min(G, npairs));
    const int grp = it / (G * nrest), r = it - grp * G * nrest;
    const int gsz = min(G, npairs - grp * G);
    pair = grp * G + r % gsz;
    rest = r / gsz;
}

template <class T, int LMAX>
struct Chunk {
    C2<T> gam[LMAX], yu[LMAX], yg[LMAX];
    C2<T> hl, ul, gl, uf, gf, hf;
};

// whether the line solvers fetch all rows of a chunk before its elimination:
// not in the register-lean variants, nor for 16-point FP64 chunks (which
// spill registers as it is); FP64 8-point chunks gain 1.2-1.4 % per step
template <class T, int LMAX, int LEAN>
constexpr bool prefetch_rows() {
    return !LEAN && (sizeof(T) == 4 || LMAX <= 8);
}

// forward elimination + first backward pass (boundary values of the chunk
// interior responses); load(i, a, b, c, d) returns row i of the interior.
// PREFETCH: all LMAX rows are fetched first (load() must accept every
// i < LMAX; rows past the chunk are not used) so that the loads overlap the
// dependent elimination chain.
template <class T, int LMAX, bool PREFETCH, class F>
__device__ __forceinline__ void chunk_eliminate(Chunk<T, LMAX> &ck, int n, F &&load) {
    using C = C2<T>;
    C binv{0, 0}, cprev{0, 0};
    constexpr int NPF = PREFETCH ? LMAX : 1;
    C pa[NPF], pb[NPF], pc[NPF], pd[NPF];
    if (PREFETCH) {
#pragma unroll
        for (int i = 0; i < LMAX; i++) load(i, pa[i], pb[i], pc[i], pd[i]);
    }
#pragma unroll
    for (int i = 0; i < LMAX; i++) {
        if (i < n) {
            C a, b, c, d;
            if (PREFETCH) {
                a = pa[i];
                b = pb[i];
                c = pc[i];
                d = pd[i];
            } else {
                load(i, a, b, c, d);
            }
            if (i == 0) {
                const C bi = crcp(b);
                ck.yu[0] = d * bi;
                ck.yg[0] = -(a * bi);
                binv = bi;
            } else {
                const C gm = cprev * binv;
                ck.gam[i - 1] = gm;
                const C bi = crcp(b - a * gm);
                ck.yu[i] = (d - a * ck.yu[i - 1]) * bi;
                ck.yg[i] = -(a * ck.yg[i - 1]) * bi;
                binv = bi;
            }
            cprev = c;
        }
    }
    ck.hl = -(cprev * binv);
    C u{0, 0}, gg{0, 0}, h{0, 0};
#pragma unroll
    for (int k = 0; k < LMAX; k++) {
        const int i = LMAX - 1 - k;
        if (i == n - 1) {
            u = ck.yu[i];
            gg = ck.yg[i];
            h = ck.hl;
            ck.ul = u;
            ck.gl = gg;
        } else if (i < n - 1) {
            u = ck.yu[i] - ck.gam[i] * u;
            gg = ck.yg[i] - ck.gam[i] * gg;
            h = -(ck.gam[i] * h);
        }
    }
    ck.uf = u;
    ck.gf = gg;
    ck.hf = h;
}

// x_i = u_i + xprev g_i + xi h_i for the chunk interior; store(i, x_i)
template <class T, int LMAX, class S>
__device__ __forceinline__ void chunk_rebuild(const Chunk<T, LMAX> &ck, int n, C2<T> xprev, C2<T> xi, S &&store) {
    C2<T> u{0, 0}, gg{0, 0}, h{0, 0};
#pragma unroll
    for (int k = 0; k < LMAX; k++) {
        const int i = LMAX - 1 - k;
        bool live = false;
        if (i == n - 1) {
            u = ck.yu[i];
            gg = ck.yg[i];
            h = ck.hl;
            live = true;
        } else if (i < n - 1) {
            u = ck.yu[i] - ck.gam[i] * u;
            gg = ck.yg[i] - ck.gam[i] * gg;
            h = -(ck.gam[i] * h);
            live = true;
        }
        if (live) store(i, u + xprev * gg + xi * h);
    }
}

template <class T>
__device__ __forceinline__ C2<T> shfl_c(C2<T> v, int lane) {
    return C2<T>{__shfl(v.x, lane), __shfl(v.y, lane)};
}

// The cyclic system of one line held by P <= 64 consecutive lanes of one wave
// (lane lb + p holds interface row p): parallel cyclic reduction through lane
// permutes, without LDS or barriers; returns xi_p.  Every lane of the wave
// calls it (lanes of inactive lines compute unused values).
template <class T>
__device__ __forceinline__ C2<T> wave_pcr(C2<T> A, C2<T> B, C2<T> Cc, C2<T> D, int P, int p, int lb) {
    using C = C2<T>;
    for (int st = 1; 2 * st < P; st <<= 1) {
        const int lm = lb + ((p - st) & (P - 1)), lq = lb + ((p + st) & (P - 1));
        const C Am = shfl_c(A, lm), Bm = shfl_c(B, lm), Cm = shfl_c(Cc, lm), Dm = shfl_c(D, lm);
        const C Aq = shfl_c(A, lq), Bq = shfl_c(B, lq), Cq = shfl_c(Cc, lq), Dq = shfl_c(D, lq);
        const C k1 = cdiv(A, Bm), k2 = cdiv(Cc, Bq);
        const C nA = -(k1 * Am), nC = -(k2 * Cq);
        B = B - k1 * Cm - k2 * Aq;
        D = D - k1 * Dm - k2 * Dq;
        A = nA;
        Cc = nC;
    }
    if (P == 1) return cdiv(D, A + B + Cc);
    const int lq = lb + ((p + P / 2) & (P - 1));
    const C Aq = shfl_c(A, lq), Bq = shfl_c(B, lq), Cq = shfl_c(Cc, lq), Dq = shfl_c(D, lq);
    const C apc = A + Cc, aqc = Aq + Cq;
    return cdiv(D * Bq - apc * Dq, B * Bq - apc * aqc);
}

// LDS of the reduced systems: first interior values of each chunk (uf, gf,
// hf) and the interface equations A xi_{p-1} + B xi_p + C xi_{p+1} = D;
// the solutions overwrite uf
template <class T>
struct RedLds {
    static constexpr int kArrays = 7;
    C2<T> *uf, *gf, *hf, *A, *B, *Cc, *D;
    __device__ __forceinline__ RedLds(C2<T> *base, int n) {
        uf = base;
        gf = uf + n;
        hf = gf + n;
        A = hf + n;
        B = A + n;
        Cc = B + n;
        D = Cc + n;
    }
};

// interface equation of chunk p from its own responses, its last row
// (ae, be, ce, de) and the first interior values of chunk p+1 at slot nx
template <class T, int LMAX>
__device__ __forceinline__ void interface_row(const Chunk<T, LMAX> &ck, C2<T> ae, C2<T> be, C2<T> ce, C2<T> de,
                                              const RedLds<T> &r, int nx, C2<T> &A, C2<T> &B, C2<T> &Cv,
                                              C2<T> &D) {
    A = ae * ck.gl;
    B = be + ae * ck.hl + ce * r.gf[nx];
    Cv = ce * r.hf[nx];
    D = de - ae * ck.ul - ce * r.uf[nx];
}

// cyclic tridiagonal system over the P chunk boundaries of one line, held in
// LDS at slots base + q*stride; every thread of the block must call it.
// Parallel cyclic reduction until the couplings reach P/2, then 2x2 pairs.
template <class T>
__device__ __forceinline__ void reduced_pcr(const RedLds<T> &r, int P, int p, int base, int stride, bool active) {
    using C = C2<T>;
    const int me = base + p * stride;
    for (int st = 1; 2 * st < P; st <<= 1) {
        C nA{0, 0}, nB{0, 0}, nC{0, 0}, nD{0, 0};
        if (active) {
            const int im = base + ((p - st) & (P - 1)) * stride, iq = base + ((p + st) & (P - 1)) * stride;
            const C k1 = cdiv(r.A[me], r.B[im]), k2 = cdiv(r.Cc[me], r.B[iq]);
            nA = -(k1 * r.A[im]);
            nC = -(k2 * r.Cc[iq]);
            nB = r.B[me] - k1 * r.Cc[im] - k2 * r.A[iq];
            nD = r.D[me] - k1 * r.D[im] - k2 * r.D[iq];
        }
        __syncthreads();
        if (active) {
            r.A[me] = nA;
            r.B[me] = nB;
            r.Cc[me] = nC;
            r.D[me] = nD;
        }
        __syncthreads();
    }
    if (active) {
        C xi;
        if (P == 1) {
            xi = cdiv(r.D[me], r.A[me] + r.B[me] + r.Cc[me]);
        } else {
            const int q = base + ((p + P / 2) & (P - 1)) * stride;
            const C apc = r.A[me] + r.Cc[me], aqc = r.A[q] + r.Cc[q];
            xi = cdiv(r.D[me] * r.B[q] - apc * r.D[q], r.B[me] * r.B[q] - apc * aqc);
        }
        r.uf[me] = xi;
    }
    __syncthreads();
}

// (row, column) of flat element e + step of rows of length M, from those of e
// (no integer division in the loops)
__device__ __forceinline__ void advance_rc(int &r, int &x, int step, int M) {
    x += step;
    while (x >= M) {
        x -= M;
        r++;
    }
}

// ---- x lines: R rows per block for both polarisations (2 R P threads,
// polarisation-major) of one component, or with nc = 2 of both components
// (4 R P threads, component-major): the two components' right-hand sides have
// different tap counts (6 and 14 for (I + mu A_y), 11 and 3 for mu O_x), so
// blocks holding both carry equal work.  The right-hand sides are staged
// through LDS in the chunk-interleaved order (element (p, i) of a row at
// i*P + p) so that global loads and stores stay coalesced; the tridiagonal
// rows come pre-interleaved from the coefficient kernel.
//   neumann 0: y = M_x (I + mu A_y) src
//   neumann 1: y = M_x (mu O_x src)
// then dst = y (skipped when dst is null) and acc = accin + y (neumann 1) or
// acc = y (neumann 0; skipped when acc is null).  accin may be src or acc
// itself: each element of accin is read and then written by the same thread.
template <class T, int LMAX, int NEU, int LEAN>
__global__ __launch_bounds__(512) void k_xsolve(const C2<T> *__restrict__ src, C2<T> *__restrict__ dst,
                                                 const C2<T> *accin, C2<T> *acc, const C2<T> *__restrict__ cf,
                                                 const C2<T> *__restrict__ xt, Geo g,
                                                 const PropParams *__restrict__ pp, LinePlan lp, int R,
                                                 int nitems_pairs, int order, int nc) {
    using C = C2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int P = lp.P, SL = LMAX * P, M = g.Mx, RP = R * P, NL = nc * 2 * R;  // NL lines per block
    C *sd = reinterpret_cast<C *>(smem_raw);  // [component][pol][row][SL]
    const RedLds<T> red(sd + NL * SL, NL * P);
    int *smap = reinterpret_cast<int *>(sd + NL * SL + RedLds<T>::kArrays * NL * P);  // x -> i*P + p
    // work items (component and row block in the order `order & 1` selects,
    // polarisation pair outermost or, with `order & 4`, innermost); each XCD
    // takes a contiguous range of them, so both components of a row block, the
    // neighbouring row blocks or the wavevectors of one wavelength meet in L2
    const int nrb = (g.My + R - 1) / R, nci = 3 - nc;  // component items per row block
    const int item = xcd_tile(blockIdx.x, gridDim.x);
    if (item >= nrb * nci * nitems_pairs) return;
    int comp, row0, pair, rest;
    if (order & 8) {  // pairs of one wavelength innermost, wavelength groups outermost
        split_grouped(item, nci * nrb, nitems_pairs, order >> 4, pair, rest);
    } else if (order & 4) {  // pair fastest: the wavevectors of one wavelength share the coefficients
        pair = item % nitems_pairs;
        rest = item / nitems_pairs;
    } else {
        pair = item / (nci * nrb);
        rest = item % (nci * nrb);
    }
    if (nc == 2) {
        comp = 0;
        row0 = rest * R;
    } else if (order & 1) {  // component before row block: both components of a row block together
        comp = rest & 1;
        row0 = (rest >> 1) * R;
    } else {  // neighbouring row blocks together
        row0 = (rest % nrb) * R;
        comp = rest / nrb;
    }
    const int prop0 = pair * 2;
    if ((int)threadIdx.x < P) {
        const int pc = threadIdx.x, s0c = chunk_start(lp, pc), lc = chunk_len(lp, pc);
        for (int i = 0; i < lc; i++) smap[s0c + i] = i * P + pc;
    }
    __syncthreads();
    const int slot = pp[prop0].slot;
    const size_t NP = (size_t)g.NP;
    const C *cfs = cf + (size_t)slot * cf_slot_size(g.Mx, g.My);
    const PairFields<T> fin(src, prop0, NP);
    const int nrow = min(R, g.My - row0), bs = blockDim.x, ntot = nrow * M;

    // a thread stores at most LMAX/2 points of the block's nc * nrow * M
    // (nrow*M <= R*P*LMAX, 2*nc*R*P threads); element e is point e % ntot of
    // component e / ntot.  The accumulator values of the Neumann launches are
    // fetched here, ahead of the gathers, so their round trip overlaps the solve.
    constexpr int NO = LMAX / 2;
    C o0v[NO], o1v[NO];
    int rr0 = (int)threadIdx.x / M, ex0 = (int)threadIdx.x - rr0 * M, cc0 = 0;
    while (rr0 >= nrow) {
        rr0 -= nrow;
        cc0++;
    }
    auto step_rc = [&](int &rr, int &x, int &cc) {
        advance_rc(rr, x, bs, M);
        while (rr >= nrow) {
            rr -= nrow;
            cc++;
        }
    };
    const size_t pbase = (size_t)prop0 * 2 * NP;
    if (NEU) {
        int rr = rr0, x = ex0, cc = cc0;
#pragma unroll
        for (int k = 0; k < NO; k++) {
            if (cc < nc) {
                const size_t o = pbase + (size_t)(nc == 2 ? cc : comp) * NP + x + M * (row0 + rr);
                o0v[k] = accin[o];
                o1v[k] = accin[o + 2 * NP];
            }
            step_rc(rr, x, cc);
        }
    }

    // right-hand sides of the block's rows, both polarisations
    for (int c = 0; c < nc; c++) {
        constexpr int KIND = NEU ? RHS_OX : RHS_YPLUS, NPR = LEAN ? 1 : rhs_points<T, LMAX>();
        C *sd0 = sd + c * 2 * R * SL, *sd1 = sd0 + R * SL;
        const int gc = nc == 2 ? c : comp;
        int r0 = (int)threadIdx.x / M, x0 = (int)threadIdx.x - r0 * M;
        for (int e0 = threadIdx.x; e0 < ntot; e0 += NPR * bs) {
            unsigned xs[NPR], ys[NPR];
            bool v[NPR];
            int slot_of[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                v[k] = e0 + k * bs < ntot;
                xs[k] = x0;
                ys[k] = row0 + r0;
                slot_of[k] = v[k] ? r0 * SL + smap[x0] : 0;
                advance_rc(r0, x0, bs, M);
            }
            C d[NPR][2];
            rhs_pair_n<T, KIND, NPR>(gc, xs, ys, v, cfs, fin, g, d);
#pragma unroll
            for (int k = 0; k < NPR; k++)
                if (v[k]) {
                    sd0[slot_of[k]] = d[k][0];
                    sd1[slot_of[k]] = d[k][1];
                }
        }
    }
    __syncthreads();

    // one chunk of one row of one polarisation (of one component) per thread
    const int ct = (int)threadIdx.x / (2 * RP), tc = (int)threadIdx.x - ct * 2 * RP;
    const int pol = tc / RP, t = tc - pol * RP, r = t / P, p = t - r * P, y = row0 + r;
    const int tcomp = nc == 2 ? ct : comp;
    const bool active = r < nrow;
    const int n = chunk_len(lp, p) - 1, line = (ct * 2 + pol) * R + r;
    C *sdl = sd + line * SL;
    const int64_t xrow = ((((int64_t)slot * 2 + tcomp) * 3) * g.My + min(y, g.My - 1)) * SL;
    const int64_t xstride = (int64_t)g.My * SL;
    const C *xa = xt + xrow, *xb = xa + xstride, *xc = xb + xstride;
    const int me = line * P + p, nx = line * P + ((p + 1) & (P - 1));
    Chunk<T, LMAX> ck;
    ck.uf = ck.gf = ck.hf = ck.ul = ck.gl = ck.hl = C{0, 0};
    if (active) {
        chunk_eliminate<T, LMAX, prefetch_rows<T, LMAX, LEAN>()>(ck, n, [&](int i, C &a, C &b, C &c, C &d) {
            const int q = i * P + p;
            a = xa[q];
            b = xb[q];
            c = xc[q];
            d = sdl[q];
        });
    }
    C xi, xprev;
    if (P <= 64) {
        // the P chunks of a line are consecutive lanes of one wave
        const int lb = ((int)threadIdx.x & 63) - p, ln = lb + ((p + 1) & (P - 1));
        const C ufn = shfl_c(ck.uf, ln), gfn = shfl_c(ck.gf, ln), hfn = shfl_c(ck.hf, ln);
        C A{0, 0}, B{0, 0}, Cv{0, 0}, D{0, 0};
        if (active) {
            const int q = n * P + p;
            const C ae = xa[q], be = xb[q], ce = xc[q];
            A = ae * ck.gl;
            B = be + ae * ck.hl + ce * gfn;
            Cv = ce * hfn;
            D = sdl[q] - ae * ck.ul - ce * ufn;
        }
        xi = wave_pcr(A, B, Cv, D, P, p, lb);
        xprev = shfl_c(xi, lb + ((p - 1) & (P - 1)));
    } else {
        if (active) {
            red.uf[me] = ck.uf;
            red.gf[me] = ck.gf;
            red.hf[me] = ck.hf;
        }
        __syncthreads();
        C A{0, 0}, B{0, 0}, Cv{0, 0}, D{0, 0};
        if (active) {
            const int q = n * P + p;
            interface_row(ck, xa[q], xb[q], xc[q], sdl[q], red, nx, A, B, Cv, D);
        }
        __syncthreads();
        if (active) {
            red.A[me] = A;
            red.B[me] = B;
            red.Cc[me] = Cv;
            red.D[me] = D;
        }
        __syncthreads();
        reduced_pcr(red, P, p, line * P, 1, active);
        xi = red.uf[me];
        xprev = red.uf[line * P + ((p - 1) & (P - 1))];
    }
    if (active) {
        chunk_rebuild(ck, n, xprev, xi, [&](int i, C v) { sdl[i * P + p] = v; });
        sdl[n * P + p] = xi;
    }
    __syncthreads();

    {
        int rr = rr0, x = ex0, cc = cc0;
#pragma unroll
        for (int k = 0; k < NO; k++) {
            if (cc < nc) {
                const int ia = (cc * 2 * R + rr) * SL + smap[x];
                const size_t o = pbase + (size_t)(nc == 2 ? cc : comp) * NP + x + M * (row0 + rr);
                const C v0 = sd[ia], v1 = sd[ia + R * SL];
                if (dst) {
                    dst[o] = v0;
                    dst[o + 2 * NP] = v1;
                }
                if (acc) {
                    acc[o] = NEU ? o0v[k] + v0 : v0;
                    acc[o + 2 * NP] = NEU ? o1v[k] + v1 : v1;
                }
            }
            step_rc(rr, x, cc);
        }
    }
}

// the DY00 (comp 0) or DY11 (comp 1) planes of point q: the tridiagonal row
// of the y-line system of that component (complex planes, selected without
// a branch so that the rows of a chunk are fetched together)
template <class T>
__device__ __forceinline__ void dy_row(int comp, const C2<T> *cfs, unsigned x, unsigned y, unsigned Mx, C2<T> &a,
                                       C2<T> &b, C2<T> &c) {
    static_assert(cf_complex(CF_DY + 0) && cf_complex(CF_DY + 1) && cf_complex(CF_DY + 2) &&
                  cf_complex(CF_DY + 9) && cf_complex(CF_DY + 10) && cf_complex(CF_DY + 11) &&
                  cf_index(CF_DY + 1) == cf_index(CF_DY + 0) + 1 && cf_index(CF_DY + 2) == cf_index(CF_DY + 0) + 2 &&
                  cf_index(CF_DY + 10) == cf_index(CF_DY + 9) + 1 && cf_index(CF_DY + 11) == cf_index(CF_DY + 9) + 2,
                  "y-line rows must be consecutive complex planes");
    constexpr int i0 = cf_index(CF_DY + 0), i1 = cf_index(CF_DY + 9);
    const C2<T> *pc = cfs + cf_point(x, y, Mx) + (comp == 0 ? i0 : i1) * kCfTile;
    a = pc[0];
    b = pc[kCfTile];
    c = pc[2 * kCfTile];
}

// right-hand sides of component `comp` (uniform per block) of one
// propagation at NPT points of one column x, rows j0, j0 + step, ... (rows
// >= M skipped); every load of every point is issued before any sum
template <class T, int KIND, int COMP, int NPT>
__device__ __forceinline__ void rhs_column_comp(unsigned x, int j0, int step, int M, const C2<T> *__restrict__ cfs,
                                                const C2<T> *__restrict__ f0, const C2<T> *__restrict__ f1,
                                                const Geo &g, C2<T> (&r)[NPT]) {
    // wide stencils (more than 8 taps) gather half of the points at a time
    constexpr int G = RhsGather<T, KIND, COMP>::N > 8 && NPT > 1 ? NPT / 2 : NPT;
#pragma unroll
    for (int k0 = 0; k0 < NPT; k0 += G) {
        RhsGather<T, KIND, COMP> gs[G];
#pragma unroll
        for (int k = 0; k < G; k++)
            if (j0 + (k0 + k) * step < M) gs[k].load(x, j0 + (k0 + k) * step, cfs, f0, f1, g, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < G; k++)
            if (j0 + (k0 + k) * step < M) r[k0 + k] = gs[k].sum();
    }
}
template <class T, int KIND, int NPT>
__device__ __forceinline__ void rhs_column(int comp, unsigned x, int j0, int step, int M,
                                           const C2<T> *__restrict__ cfs, const C2<T> *__restrict__ f0,
                                           const C2<T> *__restrict__ f1, const Geo &g, C2<T> (&r)[NPT]) {
    if (comp == 0) rhs_column_comp<T, KIND, 0, NPT>(x, j0, step, M, cfs, f0, f1, g, r);
    else rhs_column_comp<T, KIND, 1, NPT>(x, j0, step, M, cfs, f0, f1, g, r);
}

template <class T, int LMAX, int NEU, int LEAN>
__global__ __launch_bounds__(512) void k_ysolve(const C2<T> *__restrict__ src, C2<T> *__restrict__ dst,
                                                 const C2<T> *accin, C2<T> *acc, const C2<T> *__restrict__ cf, Geo g,
                                                 const PropParams *__restrict__ pp, LinePlan lp, int W,
                                                 int npairs, int order, int nc) {
    using C = C2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int P = lp.P, M = g.My, TL = LMAX * P * W;
    C *sd = reinterpret_cast<C *>(smem_raw);  // [component][TL]
    const RedLds<T> red(sd + nc * TL, nc * W * P);
    int *ymap = reinterpret_cast<int *>(sd + nc * TL + RedLds<T>::kArrays * nc * W * P);  // row j -> i*P + p
    // work items (polarisation fastest, then component and strip in the order
    // `order & 2` selects, pair outermost or, with `order & 4`, right after the
    // polarisation); each XCD takes a contiguous range, so both polarisations
    // of a strip and its components, neighbouring strips or the wavevectors of
    // one wavelength share the 128-byte lines of fields and coefficients in L2.
    // nc = 2: the block holds both components (nc*P*W threads, component-major)
    const int nstrip = (g.Mx + W - 1) / W, nci = 3 - nc;
    const int item = xcd_tile(blockIdx.x, gridDim.x);
    if (item >= nstrip * 2 * nci * npairs) return;
    const int pol = item & 1;
    int comp, lb, pair, rest;
    if (order & 8) {
        split_grouped(item >> 1, nci * nstrip, npairs, order >> 4, pair, rest);
    } else if (order & 4) {
        pair = (item >> 1) % npairs;
        rest = (item >> 1) / npairs;
    } else {
        pair = (item >> 1) / (nci * nstrip);
        rest = (item >> 1) % (nci * nstrip);
    }
    if (nc == 2) {
        comp = 0;
        lb = rest;
    } else if (order & 2) {
        comp = rest & 1;
        lb = rest >> 1;
    } else {
        lb = rest % nstrip;
        comp = rest / nstrip;
    }
    const int x0 = lb * W, prop = pair * 2 + pol;
    const int slot = pp[prop].slot;
    const size_t NP = (size_t)g.NP;
    const C *cfs = cf + (size_t)slot * cf_slot_size(g.Mx, g.My);
    const C *f0 = src + (size_t)prop * 2 * NP, *f1 = f0 + NP;
    const int wv = min(W, g.Mx - x0);
    if ((int)threadIdx.x < P) {
        const int pc = threadIdx.x, s0c = chunk_start(lp, pc), lc = chunk_len(lp, pc);
        for (int i = 0; i < lc; i++) ymap[s0c + i] = i * P + pc;
    }
    __syncthreads();
    // W is a power of two and a component has P*W threads: a thread keeps its
    // column and steps P rows at a time
    const int ct = (int)threadIdx.x / (P * W), tl = (int)threadIdx.x - ct * P * W;
    const int cc = tl & (W - 1), j0 = tl >> (__ffs(W) - 1);
    const bool col_ok = cc < wv;

    // accumulator values of the Neumann launches (a thread stores at most LMAX
    // points of its column, M <= P*LMAX), fetched ahead of the gathers so that
    // their round trip overlaps the solve
    C ov[LMAX];
    if (NEU && col_ok) {
        const size_t ob = ((size_t)prop * 2 + (nc == 2 ? ct : comp)) * NP;
#pragma unroll
        for (int k = 0; k < LMAX; k++) {
            const int j = j0 + k * P;
            if (j < M) ov[k] = accin[ob + x0 + cc + g.Mx * j];
        }
    }

    // right-hand sides of the tile: with nc = 2 every thread gathers rows of
    // both components (rows ct*P + j0, + 2P, ...), so the two components'
    // unequal tap counts are shared by all threads
    if (col_ok) {
        constexpr int KIND = NEU ? RHS_OY : RHS_XPLUS, NPT = LEAN ? 1 : rhs_points<T, LMAX>();
        const int step = nc * P;
        for (int c = 0; c < nc; c++) {
            const int gc = nc == 2 ? c : comp;
            C *sdc = sd + c * TL;
            for (int j = ct * P + j0; j < M; j += NPT * step) {
                C d[NPT];
                rhs_column<T, KIND, NPT>(gc, x0 + cc, j, step, M, cfs, f0, f1, g, d);
#pragma unroll
                for (int k = 0; k < NPT; k++)
                    if (j + k * step < M) sdc[ymap[j + k * step] * W + cc] = d[k];
            }
        }
    }
    __syncthreads();

    const int cofs = cc, p = j0, x = x0 + cofs;
    const bool active = col_ok;
    const int tcomp = nc == 2 ? ct : comp;
    sd += ct * TL;
    const int rb = ct * P * W;  // this component's reduced-system entries
    const int s = chunk_start(lp, p), n = chunk_len(lp, p) - 1;
    auto tri = [&](int j, C &a, C &b, C &c) {
        j = min(j, M - 1);  // rows past the line end are fetched but not used
        C va, vb, vc;
        dy_row<T>(tcomp, cfs, x, j, g.Mx, va, vb, vc);
        a = C{-va.x, -va.y};
        b = C{(T)1 - vb.x, -vb.y};
        c = C{-vc.x, -vc.y};
    };
    const int me = rb + p * W + cofs, nx = rb + ((p + 1) & (P - 1)) * W + cofs;
    Chunk<T, LMAX> ck;
    if (active) {
        chunk_eliminate<T, LMAX, prefetch_rows<T, LMAX, LEAN>()>(ck, n, [&](int i, C &a, C &b, C &c, C &d) {
            tri(s + i, a, b, c);
            d = sd[(i * P + p) * W + cofs];
        });
        red.uf[me] = ck.uf;
        red.gf[me] = ck.gf;
        red.hf[me] = ck.hf;
    }
    __syncthreads();
    C A{0, 0}, B{0, 0}, Cv{0, 0}, D{0, 0};
    if (active) {
        C ae, be, ce;
        tri(s + n, ae, be, ce);
        interface_row(ck, ae, be, ce, sd[(n * P + p) * W + cofs], red, nx, A, B, Cv, D);
    }
    __syncthreads();
    if (active) {
        red.A[me] = A;
        red.B[me] = B;
        red.Cc[me] = Cv;
        red.D[me] = D;
    }
    __syncthreads();
    reduced_pcr(red, P, p, rb + cofs, W, active);
    if (active) {
        const C xi = red.uf[me], xprev = red.uf[rb + ((p - 1) & (P - 1)) * W + cofs];
        chunk_rebuild(ck, n, xprev, xi, [&](int i, C v) { sd[(i * P + p) * W + cofs] = v; });
        sd[(n * P + p) * W + cofs] = xi;
    }
    __syncthreads();

    {
        // a thread stores at most LMAX points of its column (M <= P*LMAX)
        const size_t ob = ((size_t)prop * 2 + tcomp) * NP;
        if (col_ok) {
#pragma unroll
            for (int k = 0; k < LMAX; k++) {
                const int j = j0 + k * P;
                if (j < M) {
                    const C v = sd[ymap[j] * W + cc];
                    const size_t q = ob + x0 + cc + g.Mx * j;
                    if (dst) dst[q] = v;
                    if (acc) acc[q] = NEU ? ov[k] + v : v;
                }
            }
        }
    }
}

// ---- y lines, both polarisations per block (2 W P threads,
// polarisation-major): the right-hand sides of the W x M tile are gathered for
// the two polarisations together (one set of coefficient loads), each chunk
// thread then solves one polarisation as in k_ysolve.
template <class T, int LMAX, int NEU, int LEAN>
__global__ __launch_bounds__(512) void k_ysolve_pair(const C2<T> *__restrict__ src, C2<T> *__restrict__ dst,
                                                      const C2<T> *accin, C2<T> *acc,
                                                      const C2<T> *__restrict__ cf, Geo g,
                                                      const PropParams *__restrict__ pp, LinePlan lp, int W,
                                                      int npairs, int order) {
    using C = C2<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int P = lp.P, M = g.My, TL = LMAX * P * W, WP = W * P;
    C *sd = reinterpret_cast<C *>(smem_raw);  // [pol][TL]
    const RedLds<T> red(sd + 2 * TL, 2 * WP);
    int *ymap = reinterpret_cast<int *>(sd + 2 * TL + RedLds<T>::kArrays * 2 * WP);  // row j -> i*P + p
    // work items (component and strip in the order `order & 2` selects, pair
    // outermost or, with `order & 4`, innermost), XCD-contiguous ranges
    const int nstrip = (g.Mx + W - 1) / W;
    const int item = xcd_tile(blockIdx.x, gridDim.x);
    if (item >= nstrip * 2 * npairs) return;
    int comp, lb, pair, rest;
    if (order & 8) {
        split_grouped(item, 2 * nstrip, npairs, order >> 4, pair, rest);
    } else if (order & 4) {
        pair = item % npairs;
        rest = item / npairs;
    } else {
        pair = item / (2 * nstrip);
        rest = item % (2 * nstrip);
    }
    if (order & 2) {
        comp = rest & 1;
        lb = rest >> 1;
    } else {
        lb = rest % nstrip;
        comp = rest / nstrip;
    }
    const int x0 = lb * W, prop0 = pair * 2;
    const int slot = pp[prop0].slot;
    const size_t NP = (size_t)g.NP;
    const C *cfs = cf + (size_t)slot * cf_slot_size(g.Mx, g.My);
    const PairFields<T> fin(src, prop0, NP);
    const int wv = min(W, g.Mx - x0), lw = __ffs(W) - 1;
    if ((int)threadIdx.x < P) {
        const int pc = threadIdx.x, s0c = chunk_start(lp, pc), lc = chunk_len(lp, pc);
        for (int i = 0; i < lc; i++) ymap[s0c + i] = i * P + pc;
    }
    __syncthreads();
    C *sd0 = sd, *sd1 = sd + TL;
    // staging and output: a thread keeps column cc and steps 2P rows at a time
    const int cc = threadIdx.x & (W - 1), jr = (int)threadIdx.x >> lw;
    const bool col_ok = cc < wv;
    // accumulator values of the Neumann launches (at most LMAX/2 rows of each
    // polarisation per thread, rows step 2P), fetched ahead of the gathers
    const size_t o0 = ((size_t)prop0 * 2 + comp) * NP, o1 = o0 + 2 * NP;
    constexpr int NO = (LMAX + 1) / 2;
    C o0v[NO], o1v[NO];
    if (NEU && col_ok) {
#pragma unroll
        for (int k = 0; k < NO; k++) {
            const int j = jr + k * 2 * P;
            if (j < M) {
                const unsigned q = x0 + cc + g.Mx * j;
                o0v[k] = accin[o0 + q];
                o1v[k] = accin[o1 + q];
            }
        }
    }

    if (col_ok) {
        constexpr int KIND = NEU ? RHS_OY : RHS_XPLUS, NPR = LEAN ? 1 : rhs_points<T, LMAX>();
        for (int j = jr; j < M; j += NPR * 2 * P) {
            unsigned xs[NPR], ys[NPR];
            bool v[NPR];
#pragma unroll
            for (int k = 0; k < NPR; k++) {
                v[k] = j + k * 2 * P < M;
                xs[k] = x0 + cc;
                ys[k] = v[k] ? j + k * 2 * P : 0;
            }
            C d[NPR][2];
            rhs_pair_n<T, KIND, NPR>(comp, xs, ys, v, cfs, fin, g, d);
#pragma unroll
            for (int k = 0; k < NPR; k++)
                if (v[k]) {
                    const int ia = ymap[ys[k]] * W + cc;
                    sd0[ia] = d[k][0];
                    sd1[ia] = d[k][1];
                }
        }
    }
    __syncthreads();

    // one chunk of one column of one polarisation per thread
    const int pol = threadIdx.x / WP, t = threadIdx.x - pol * WP, cofs = t & (W - 1), p = t >> lw, x = x0 + cofs;
    const bool active = cofs < wv;
    C *sdl = sd + pol * TL;
    const int s = chunk_start(lp, p), n = chunk_len(lp, p) - 1;
    auto tri = [&](int j, C &a, C &b, C &c) {
        j = min(j, M - 1);  // rows past the line end are fetched but not used
        C va, vb, vc;
        dy_row<T>(comp, cfs, x, j, g.Mx, va, vb, vc);
        a = C{-va.x, -va.y};
        b = C{(T)1 - vb.x, -vb.y};
        c = C{-vc.x, -vc.y};
    };
    const int base = pol * WP + cofs, me = base + p * W, nx = base + ((p + 1) & (P - 1)) * W;
    Chunk<T, LMAX> ck;
    if (active) {
        chunk_eliminate<T, LMAX, prefetch_rows<T, LMAX, LEAN>()>(ck, n, [&](int i, C &a, C &b, C &c, C &d) {
            tri(s + i, a, b, c);
            d = sdl[(i * P + p) * W + cofs];
        });
        red.uf[me] = ck.uf;
        red.gf[me] = ck.gf;
        red.hf[me] = ck.hf;
    }
    __syncthreads();
    C A{0, 0}, B{0, 0}, Cv{0, 0}, D{0, 0};
    if (active) {
        C ae, be, ce;
        tri(s + n, ae, be, ce);
        interface_row(ck, ae, be, ce, sdl[(n * P + p) * W + cofs], red, nx, A, B, Cv, D);
    }
    __syncthreads();
    if (active) {
        red.A[me] = A;
        red.B[me] = B;
        red.Cc[me] = Cv;
        red.D[me] = D;
    }
    __syncthreads();
    reduced_pcr(red, P, p, base, W, active);
    if (active) {
        const C xi = red.uf[me], xprev = red.uf[base + ((p - 1) & (P - 1)) * W];
        chunk_rebuild(ck, n, xprev, xi, [&](int i, C v) { sdl[(i * P + p) * W + cofs] = v; });
        sdl[(n * P + p) * W + cofs] = xi;
    }
    __syncthreads();

    {
        // a thread stores at most LMAX/2 points of its column for each
        // polarisation (M <= P*LMAX, rows step 2P); accumulators in one round trip
        if (col_ok) {
#pragma unroll
            for (int k = 0; k < NO; k++) {
                const int j = jr + k * 2 * P;
                if (j < M) {
                    const int ia = ymap[j] * W + cc;
                    const C v0 = sd0[ia], v1 = sd1[ia];
                    const unsigned q = x0 + cc + g.Mx * j;
                    if (dst) {
                        dst[o0 + q] = v0;
                        dst[o1 + q] = v1;
                    }
                    if (acc) {
                        acc[o0 + q] = NEU ? o0v[k] + v0 : v0;
                        acc[o1 + q] = NEU ? o1v[k] + v1 : v1;
                    }
                }
            }
        }
    }
}

// dynamic LDS above the 64 KiB default needs an opt-in per kernel
size_t max_block_lds() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 64 * 1024;
    return (size_t)v;
}

void allow_kernel_lds(const void *f, size_t lds) {
    if (lds <= 64 * 1024) return;
    static std::mutex mu;
    static std::map<const void *, size_t> granted;
    std::lock_guard<std::mutex> lk(mu);
    auto it = granted.find(f);
    if (it != granted.end() && it->second >= lds) return;
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        throw std::runtime_error("the device refuses " + std::to_string(lds) + " bytes of LDS per block");
    granted[f] = lds;
}

template <class K>
static void allow_lds(K kernel, size_t lds) {
    allow_kernel_lds(reinterpret_cast<const void *>(kernel), lds);
}

LinePlan make_line_plan(int M, int chunks) {
    LinePlan lp{};
    lp.M = M;
    int cap = std::max(1, M / 2);
    int P = 1;
    while (P * 2 <= std::min(cap, chunks)) P *= 2;
    lp.P = P;
    lp.base = M / P;
    lp.rem = M % P;
    int need = lp.base + (lp.rem ? 1 : 0);
    lp.lmax = need <= 4 ? 4 : need <= 8 ? 8 : need <= 16 ? 16 : 0;
    if (lp.lmax == 0) throw std::runtime_error("line solver: chunks longer than 16 points (too few chunks)");
    return lp;
}

// staged right-hand sides of both polarisations + their reduced systems
template <class T>
size_t xsolve_lds_bytes(const LinePlan &lp, int R, int nc) {
    return (size_t)nc * 2 * R * ((size_t)lp.lmax * lp.P + (size_t)RedLds<T>::kArrays * lp.P) * sizeof(C2<T>) +
           (size_t)lp.M * sizeof(int);
}
template <class T>
size_t ysolve_lds_bytes(const LinePlan &lp, int W, int pair) {
    // pair: 1 both polarisations, 2 both components (one polarisation)
    return (size_t)(pair ? 2 : 1) * W * ((size_t)lp.lmax * lp.P + (size_t)RedLds<T>::kArrays * lp.P) * sizeof(C2<T>) +
           (size_t)lp.M * sizeof(int);
}

template <class T, int NEU, int LEAN>
static void xsolve_dispatch(const LinePlan &lp, int R, dim3 grid, const C2<T> *src, C2<T> *dst, const C2<T> *accin,
                            C2<T> *acc,
                            const C2<T> *cf, const C2<T> *xt, const Geo &g, const PropParams *pp, int npairs,
                            int order, int nc, hipStream_t st) {
    const dim3 block(nc * 2 * lp.P * R);
    const size_t lds = xsolve_lds_bytes<T>(lp, R, nc);
    switch (lp.lmax) {
        case 4:
            allow_lds(k_xsolve<T, 4, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_xsolve<T, 4, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, xt, g, pp, lp, R, npairs, order, nc);
            break;
        case 8:
            allow_lds(k_xsolve<T, 8, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_xsolve<T, 8, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, xt, g, pp, lp, R, npairs, order, nc);
            break;
        case 16:
            allow_lds(k_xsolve<T, 16, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_xsolve<T, 16, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, xt, g, pp, lp, R, npairs, order, nc);
            break;
        default: throw std::runtime_error("x line solver: unsupported chunk length");
    }
    NMK_LAUNCH_CHECK();
}

template <class T, int NEU, int LEAN>
static void ysolve_pair_dispatch(const LinePlan &lp, int W, dim3 grid, const C2<T> *src, C2<T> *dst, const C2<T> *accin,
                            C2<T> *acc,
                                 const C2<T> *cf, const Geo &g, const PropParams *pp, int npairs, int order,
                                 hipStream_t st) {
    const dim3 block(2 * lp.P * W);
    const size_t lds = ysolve_lds_bytes<T>(lp, W, 1);
    switch (lp.lmax) {
        case 4:
            allow_lds(k_ysolve_pair<T, 4, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_ysolve_pair<T, 4, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, g, pp, lp, W,
                               npairs, order);
            break;
        case 8:
            allow_lds(k_ysolve_pair<T, 8, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_ysolve_pair<T, 8, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, g, pp, lp, W,
                               npairs, order);
            break;
        case 16:
            allow_lds(k_ysolve_pair<T, 16, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_ysolve_pair<T, 16, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, g, pp, lp,
                               W, npairs, order);
            break;
        default: throw std::runtime_error("y line solver: unsupported chunk length");
    }
    NMK_LAUNCH_CHECK();
}

template <class T, int NEU, int LEAN>
static void ysolve_dispatch(const LinePlan &lp, int W, dim3 grid, const C2<T> *src, C2<T> *dst, const C2<T> *accin,
                            C2<T> *acc,
                            const C2<T> *cf, const Geo &g, const PropParams *pp, int npairs, int order, int nc,
                            hipStream_t st) {
    const dim3 block(nc * lp.P * W);
    const size_t lds = ysolve_lds_bytes<T>(lp, W, nc == 2 ? 2 : 0);
    switch (lp.lmax) {
        case 4:
            allow_lds(k_ysolve<T, 4, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_ysolve<T, 4, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, g, pp, lp, W, npairs, order, nc);
            break;
        case 8:
            allow_lds(k_ysolve<T, 8, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_ysolve<T, 8, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, g, pp, lp, W, npairs, order, nc);
            break;
        case 16:
            allow_lds(k_ysolve<T, 16, NEU, LEAN>, lds);
            hipLaunchKernelGGL((k_ysolve<T, 16, NEU, LEAN>), grid, block, lds, st, src, dst, accin, acc, cf, g, pp, lp, W, npairs, order, nc);
            break;
        default: throw std::runtime_error("y line solver: unsupported chunk length");
    }
    NMK_LAUNCH_CHECK();
}

template <class T>
void launch_xsolve(int neumann, const C2<T> *src, C2<T> *dst, const C2<T> *accin, C2<T> *acc, const C2<T> *cf,
                   const C2<T> *xt,
                   const Geo &g, const PropParams *pp, int nprops, const LinePlan &lp, int R, int order,
                   int lean, int both, hipStream_t st) {
    const int nc = both ? 2 : 1;
    if (nc * 2 * lp.P * R > 512 || lp.M != g.Mx || R < 1 || lp.P * lp.lmax < lp.M)
        throw std::runtime_error("x line solver: bad launch shape");
    if (nprops % 2) throw std::runtime_error("x line solver: propagations come in polarisation pairs");
    const int npairs = nprops / 2;
    const int64_t items = (int64_t)((g.My + R - 1) / R) * (3 - nc) * npairs;
    dim3 grid((unsigned)((items + 7) / 8 * 8));
    if (lean) {
        if (neumann) xsolve_dispatch<T, 1, 1>(lp, R, grid, src, dst, accin, acc, cf, xt, g, pp, npairs, order, nc, st);
        else xsolve_dispatch<T, 0, 1>(lp, R, grid, src, dst, accin, acc, cf, xt, g, pp, npairs, order, nc, st);
    } else {
        if (neumann) xsolve_dispatch<T, 1, 0>(lp, R, grid, src, dst, accin, acc, cf, xt, g, pp, npairs, order, nc, st);
        else xsolve_dispatch<T, 0, 0>(lp, R, grid, src, dst, accin, acc, cf, xt, g, pp, npairs, order, nc, st);
    }
}

template <class T>
void launch_ysolve(int neumann, const C2<T> *src, C2<T> *dst, const C2<T> *accin, C2<T> *acc, const C2<T> *cf,
                   const Geo &g,
                   const PropParams *pp, int nprops, const LinePlan &lp, int W, int order, int lean, int pair,
                   hipStream_t st) {
    if ((pair ? 2 : 1) * lp.P * W > 512 || lp.M != g.My || W < 1 || (W & (W - 1)) || lp.P * lp.lmax < lp.M ||
        pair < 0 || pair > 2)
        throw std::runtime_error("y line solver: bad launch shape");
    if (nprops % 2) throw std::runtime_error("y line solver: propagations come in polarisation pairs");
    const int npairs = nprops / 2;
    if (pair == 1) {
        const int64_t items = (int64_t)((g.Mx + W - 1) / W) * 2 * npairs;
        dim3 grid((unsigned)((items + 7) / 8 * 8));
        if (lean) {
            if (neumann) ysolve_pair_dispatch<T, 1, 1>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, st);
            else ysolve_pair_dispatch<T, 0, 1>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, st);
        } else {
            if (neumann) ysolve_pair_dispatch<T, 1, 0>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, st);
            else ysolve_pair_dispatch<T, 0, 0>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, st);
        }
        return;
    }
    const int nc = pair == 2 ? 2 : 1;
    const int64_t items = (int64_t)((g.Mx + W - 1) / W) * 2 * (3 - nc) * npairs;
    dim3 grid((unsigned)((items + 7) / 8 * 8));
    if (lean) {
        if (neumann) ysolve_dispatch<T, 1, 1>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, nc, st);
        else ysolve_dispatch<T, 0, 1>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, nc, st);
    } else {
        if (neumann) ysolve_dispatch<T, 1, 0>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, nc, st);
        else ysolve_dispatch<T, 0, 0>(lp, W, grid, src, dst, accin, acc, cf, g, pp, npairs, order, nc, st);
    }
}

// ---------------------------------------------------------------------------
// output: march grid -> full grid (duplicates mirror column/row 0), global
// shift phase, output Fresnel with each point's own eps at the last plane
// (BeamPropagationLib/FresnelOperator.cpp:45-69)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_output(const C2<T> *__restrict__ F, const double *__restrict__ lc,
                                                const uint8_t *__restrict__ mask, Geo g,
                                                const SlotParams *__restrict__ sp, const PropParams *__restrict__ pp,
                                                c64 *__restrict__ E) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.Nfull) return;
    const int prop = blockIdx.y;
    const PropParams p = pp[prop];
    const SlotParams s = sp[p.slot];
    const int ix = (int)(m % g.Nx), iy = (int)(m / g.Nx);
    const int jx = g.px ? ix % g.Mx : ix, jy = g.py ? iy % g.My : iy;
    const int64_t ip = jx + (int64_t)g.Mx * jy;
    const c64 ph = mk(p.ph_re, p.ph_im);
    const c64 a = ph * cast2<double>(F[((int64_t)prop * 2) * g.NP + ip]);
    const c64 b = ph * cast2<double>(F[((int64_t)prop * 2 + 1) * g.NP + ip]);
    double e[12];
    eps_voxel(lc, mask, g.field_dim, m + g.Nfull * (g.Nz - 1), s, e);
    const double n = s.nout_fresnel, nxx = e[9], nyy = e[10], nxy = e[11];
    const double det = (n + nxx) * (n + nyy) - nxy * nxy;
    const double cxx = 1. - n * (n + nyy) / det, cyy = 1. - n * (n + nxx) / det, cxy = n * nxy / det;
    E[((int64_t)2 * prop) * g.Nfull + m] = 2.0 * (cxx * a + cxy * b);
    E[((int64_t)2 * prop + 1) * g.Nfull + m] = 2.0 * (cyy * b + cxy * a);
}

template <class T>
void launch_output(const C2<T> *F, const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                   const PropParams *pp, int nprops, int nq, c64 *E, hipStream_t st) {
    (void)nq;
    hipLaunchKernelGGL(k_output<T>, dim3(blocks_for(g.Nfull, 256), nprops), dim3(256), 0, st, F, lc, mask, g, sp, pp, E);
    NMK_LAUNCH_CHECK();
}

// Output of a mesh of one plane (no slab marched, BPMIteration.cpp:127): every
// full-grid point, duplicate column and row included, keeps its own entrance
// field (the beam through the input interface with that point's
// permittivity, or for beam 2 the caller's fields, read in place from E) and
// goes through the output interface (FresnelOperator.cpp:45-69).
__global__ __launch_bounds__(256) void k_output_entrance(const double *__restrict__ lc,
                                                         const uint8_t *__restrict__ mask, Geo g,
                                                         const SlotParams *__restrict__ sp,
                                                         const PropParams *__restrict__ pp, int beam, double waist,
                                                         double dx, double dy, c64 *__restrict__ E) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.Nfull) return;
    const int prop = blockIdx.y;
    const PropParams p = pp[prop];
    const SlotParams s = sp[p.slot];
    c64 *e0 = E + ((int64_t)2 * prop) * g.Nfull, *e1 = e0 + g.Nfull;
    c64 a, b;
    if (beam == 2) {
        a = e0[m];
        b = e1[m];
    } else {
        double ar, br;
        entrance_field(lc, mask, g, s, p, (int)(m % g.Nx), (int)(m / g.Nx), beam == 1, waist, dx, dy, ar, br);
        a = mk(ar, 0.0);
        b = mk(br, 0.0);
    }
    const c64 ph = mk(p.ph_re, p.ph_im);
    a = ph * a;
    b = ph * b;
    double e[12];
    eps_voxel(lc, mask, g.field_dim, m + g.Nfull * (g.Nz - 1), s, e);
    const double n = s.nout_fresnel, nxx = e[9], nyy = e[10], nxy = e[11];
    const double det = (n + nxx) * (n + nyy) - nxy * nxy;
    const double cxx = 1. - n * (n + nyy) / det, cyy = 1. - n * (n + nxx) / det, cxy = n * nxy / det;
    e0[m] = 2.0 * (cxx * a + cxy * b);
    e1[m] = 2.0 * (cyy * b + cxy * a);
}

void launch_output_entrance(const double *lc, const uint8_t *mask, const Geo &g, const SlotParams *sp,
                            const PropParams *pp, int nprops, int beam, double waist, double dx, double dy, c64 *E,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_output_entrance, dim3(blocks_for(g.Nfull, 256), nprops), dim3(256), 0, st, lc, mask, g, sp,
                       pp, beam, waist, dx, dy, E);
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// one transverse plane of the bulk output (BPMIteration.cpp:111-120, 153-168):
// march state -> full grid (duplicate column/row copied) with the phase of the
// 2*plane shift half-steps applied so far, no interface;
// bulk[(((iwl*nq + iq)*2 + pol)*Nz + plane)*2 + comp][m]
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_bulk_plane(const C2<T> *__restrict__ F, Geo g,
                                                    const SlotParams *__restrict__ sp,
                                                    const PropParams *__restrict__ pp, int plane, int nq,
                                                    c64 *__restrict__ bulk) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.Nfull) return;
    const int prop = blockIdx.y;
    const PropParams p = pp[prop];
    const int iwl = sp[p.slot].iwl;
    const int ix = (int)(m % g.Nx), iy = (int)(m / g.Nx);
    const int jx = g.px ? ix % g.Mx : ix, jy = g.py ? iy % g.My : iy;
    const int64_t ip = jx + (int64_t)g.Mx * jy;
    double s, c;
    sincos(p.ph_half * 2.0 * plane, &s, &c);
    const c64 ph = mk(c, s);
    const c64 a = ph * cast2<double>(F[((int64_t)prop * 2) * g.NP + ip]);
    const c64 b = ph * cast2<double>(F[((int64_t)prop * 2 + 1) * g.NP + ip]);
    c64 *o = bulk + ((((int64_t)iwl * nq + p.iq) * 2 + p.pol) * g.Nz + plane) * 2 * g.Nfull;
    o[m] = a;
    o[g.Nfull + m] = b;
}

template <class T>
void launch_bulk_plane(const C2<T> *F, const Geo &g, const SlotParams *sp, const PropParams *pp, int nprops,
                       int plane, int nq, c64 *bulk, hipStream_t st) {
    hipLaunchKernelGGL(k_bulk_plane<T>, dim3(blocks_for(g.Nfull, 256), nprops), dim3(256), 0, st, F, g, sp, pp,
                       plane, nq, bulk);
    NMK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// screen filter FFT: rows as one line per workgroup, columns as C adjacent
// columns per workgroup (rows of C consecutive values in HBM), both with the
// Stockham stages of fft_device.h
// ---------------------------------------------------------------------------
bool fft_supported(int n) {
    if (n < 1) return false;
    int m = n;
    for (int r : {8, 4, 2, 3, 5, 7, 11, 13})
        while (m % r == 0) m /= r;
    return m == 1;
}

FftPlan make_fft_plan(int n) {
    FftPlan pl{};
    pl.n = n;
    if (n < 1) throw std::runtime_error("FFT length must be positive");
    if (!fft_supported(n) || n > 2048) {
        pl.direct = 1;
        return pl;
    }
    int m = n;
    pl.nst = 0;
    for (int r : {8, 4, 2, 3, 5, 7, 11, 13})
        while (m % r == 0 && m > 1) {
            if (pl.nst >= 16) throw std::runtime_error("FFT plan too deep");
            pl.radices |= (uint64_t)r << (4 * pl.nst++);
            m /= r;
        }
    if (m != 1) throw std::runtime_error("FFT length " + std::to_string(n) + " has a prime factor above 13");
    return pl;
}

int fft_cols_per_block(int n) { return fft_lines_per_block(n, 8); }

// rows: block b transforms lines C*b .. C*b+C-1 of the whole batch (each a
// contiguous run of n values); src and dst may be the same buffer
template <int C, bool ODD>
__global__ __launch_bounds__(256) void k_fft_rows(const c64 *src, c64 *dst, int64_t nlines, FftPlan pl,
                                                  const c64 *__restrict__ tw, int sign) {
    extern __shared__ c64 fbuf[];
    const int n = pl.n;
    c64 *stw = fbuf + 2 * C * n;
    const int64_t l0 = (int64_t)blockIdx.x * C;
    const int nl = (int)(nlines - l0 < C ? nlines - l0 : C);
    const int64_t base = l0 * n;
    for (int t = threadIdx.x; t < C * n; t += blockDim.x) fbuf[t] = t < nl * n ? src[base + t] : mk(0.0, 0.0);
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    const c64 *res = fft_stages<C, true, ODD>(fbuf, fbuf + C * n, pl, stw, sign);
    for (int t = threadIdx.x; t < nl * n; t += blockDim.x) dst[base + t] = res[t];
}

// columns, in place: block (group g, plane) transforms columns C*g .. C*g+C-1
template <int C, bool ODD>
__global__ __launch_bounds__(256) void k_fft_cols(c64 *__restrict__ data, int Nx, int ngroups, int64_t nblocks,
                                                  FftPlan pl, const c64 *__restrict__ tw, int sign) {
    extern __shared__ c64 fbuf[];
    const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
    if (blk >= nblocks) return;
    const int n = pl.n;
    c64 *stw = fbuf + 2 * C * n;
    const int g = (int)(blk % ngroups);
    c64 *plane = data + (blk / ngroups) * (int64_t)Nx * n;
    const int x0 = g * C;
    for (int t = threadIdx.x; t < n * C; t += blockDim.x) {
        const int i = t % C, y = t / C;
        fbuf[t] = (x0 + i < Nx) ? plane[(int64_t)y * Nx + x0 + i] : mk(0.0, 0.0);
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    const c64 *res = fft_stages<C, false, ODD>(fbuf, fbuf + n * C, pl, stw, sign);
    for (int t = threadIdx.x; t < n * C; t += blockDim.x) {
        const int i = t % C, y = t / C;
        if (x0 + i < Nx) plane[(int64_t)y * Nx + x0 + i] = res[t];
    }
}

// Direct DFT of lines (lengths the LDS plans do not take): a block holds C
// lines and the twiddle table in LDS; output k of a line is
// sum_j x_j tw[(j k) mod n] (conjugated twiddles for sign +1), with the twiddle
// index advanced by k per term (no products or divisions in the loop).
// Rows (elem_stride 1) run k fastest over one line per block, columns
// (line_stride 1) the C adjacent lines fastest, so loads and stores coalesce.
__global__ __launch_bounds__(256) void k_dft_lines(const c64 *src, c64 *dst, int n, int lines, int C,
                                                   int groups, int64_t plane_size, int64_t line_stride,
                                                   int64_t elem_stride, int sign, const c64 *__restrict__ tw) {
    extern __shared__ c64 dbuf[];
    c64 *stw = dbuf + (int64_t)C * n;
    const int64_t plane = blockIdx.x / groups;
    const int l0 = (int)(blockIdx.x % groups) * C;
    const int nl = min(C, lines - l0);
    const bool kfast = elem_stride == 1;
    const int64_t base = plane * plane_size + (int64_t)l0 * line_stride;
    const int tot = nl * n;
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        const int i = kfast ? t / n : t % nl, k = kfast ? t % n : t / nl;
        dbuf[i * n + k] = src[base + i * line_stride + k * elem_stride];
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = sign > 0 ? conj(tw[t]) : tw[t];
    __syncthreads();
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        const int i = kfast ? t / n : t % nl, k = kfast ? t % n : t / nl;
        const c64 *x = dbuf + i * n;
        c64 acc = mk(0.0, 0.0);
        int idx = 0;
        for (int j = 0; j < n; j++) {
            acc = cmad(acc, x[j], stw[idx]);
            idx += k;
            if (idx >= n) idx -= n;
        }
        dst[base + i * line_stride + k * elem_stride] = acc;
    }
}

void launch_dft_lines(const c64 *src, c64 *dst, int n, int lines, int planes, int64_t plane_size,
                      int64_t line_stride, int64_t elem_stride, int sign, const c64 *tw, hipStream_t st) {
    if (n < 1 || lines < 1 || planes < 1) return;
    // columns: up to 8 adjacent lines per block (whole 128-byte rows); rows: one
    int C = elem_stride == 1 ? 1 : 8;
    while (C > 1 && (size_t)(C + 1) * n * sizeof(c64) > 64 * 1024) C /= 2;
    const size_t lds = (size_t)(C + 1) * n * sizeof(c64);
    allow_kernel_lds(reinterpret_cast<const void *>(k_dft_lines), lds);
    const int groups = (lines + C - 1) / C;
    const int64_t blocks = (int64_t)groups * planes;
    if (blocks > 0x7fffffff) throw std::runtime_error("DFT: too many lines");
    hipLaunchKernelGGL(k_dft_lines, dim3((unsigned)blocks), dim3(256), lds, st, src, dst, n, lines, C, groups,
                       plane_size, line_stride, elem_stride, sign, tw);
    NMK_LAUNCH_CHECK();
}

void launch_fft_rows(const c64 *src, c64 *dst, int Nx, int Ny, int planes, int sign, const c64 *tw_x,
                     const FftPlan &px, hipStream_t st) {
    if (px.direct) {
        launch_dft_lines(src, dst, Nx, Ny, planes, (int64_t)Nx * Ny, Nx, 1, sign, tw_x, st);
        return;
    }
    const int C = fft_lines_per_block(Nx, 16);
    const int64_t nlines = (int64_t)Ny * planes;
    const unsigned grid = (unsigned)((nlines + C - 1) / C);
    const size_t lds = fft_lds_bytes(Nx, C);
    switch (C) {
#define NMK_ROWS(CC)                                                                                          \
    case CC:                                                                                                  \
        if (fft_pow2(px)) {                                                                                   \
            allow_kernel_lds(reinterpret_cast<const void *>(k_fft_rows<CC, false>), lds);                    \
            hipLaunchKernelGGL((k_fft_rows<CC, false>), dim3(grid), dim3(256), lds, st, src, dst, nlines, px, \
                               tw_x, sign);                                                                   \
        } else {                                                                                              \
            allow_kernel_lds(reinterpret_cast<const void *>(k_fft_rows<CC, true>), lds);                     \
            hipLaunchKernelGGL((k_fft_rows<CC, true>), dim3(grid), dim3(256), lds, st, src, dst, nlines, px,  \
                               tw_x, sign);                                                                   \
        }                                                                                                     \
        break;
        NMK_ROWS(16)
        NMK_ROWS(8)
        NMK_ROWS(4)
        NMK_ROWS(2)
        NMK_ROWS(1)
#undef NMK_ROWS
    default: throw std::runtime_error("FFT rows: bad block shape");
    }
    NMK_LAUNCH_CHECK();
}

void launch_fft_cols(c64 *data, int Nx, int Ny, int planes, int sign, const c64 *tw_y, const FftPlan &py,
                     hipStream_t st) {
    if (py.direct) {
        launch_dft_lines(data, data, Ny, Nx, planes, (int64_t)Nx * Ny, 1, Nx, sign, tw_y, st);
        return;
    }
    const int C = fft_cols_per_block(Ny);
    const int ngroups = (Nx + C - 1) / C;
    const int64_t nblocks = (int64_t)ngroups * planes;
    const unsigned grid = (unsigned)((nblocks + 7) / 8 * 8);
    const size_t lds = fft_lds_bytes(Ny, C);
    switch (C) {
#define NMK_COLS(CC)                                                                                          \
    case CC:                                                                                                  \
        if (fft_pow2(py)) {                                                                                   \
            allow_kernel_lds(reinterpret_cast<const void *>(k_fft_cols<CC, false>), lds);                    \
            hipLaunchKernelGGL((k_fft_cols<CC, false>), dim3(grid), dim3(256), lds, st, data, Nx, ngroups,    \
                               nblocks, py, tw_y, sign);                                                      \
        } else {                                                                                              \
            allow_kernel_lds(reinterpret_cast<const void *>(k_fft_cols<CC, true>), lds);                     \
            hipLaunchKernelGGL((k_fft_cols<CC, true>), dim3(grid), dim3(256), lds, st, data, Nx, ngroups,     \
                               nblocks, py, tw_y, sign);                                                      \
        }                                                                                                     \
        break;
        NMK_COLS(8)
        NMK_COLS(4)
        NMK_COLS(2)
        NMK_COLS(1)
#undef NMK_COLS
    default: throw std::runtime_error("FFT columns: bad block shape");
    }
    NMK_LAUNCH_CHECK();
}

void launch_fft2(c64 *data, int Nx, int Ny, int planes, int sign, const c64 *tw_x, const c64 *tw_y,
                 const FftPlan &px, const FftPlan &py, hipStream_t st) {
    if (Nx > 1) launch_fft_rows(data, data, Nx, Ny, planes, sign, tw_x, px, st);
    if (Ny > 1) launch_fft_cols(data, Nx, Ny, planes, sign, tw_y, py, st);
}

// ---------------------------------------------------------------------------
// per-k 2x2 transfer filter (PostprocessorLib/ScreenOutput.cpp:226-302) and
// the 1/(Nx*Ny) normalisation, applied to both polarisations (:206-219);
// any number of isotropic layers (the transfer product is accumulated in
// registers, one layer at a time)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_filter(c64 *__restrict__ E, Geo g, const SlotParams *__restrict__ sp,
                                                const double *__restrict__ qxy, int nq,
                                                const double *__restrict__ iso_thick,
                                                const double *__restrict__ iso_index, int n_iso, double dx, double dy,
                                                double z_foc, double NA) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.Nfull) return;
    const int slot = blockIdx.y / nq, iq = blockIdx.y % nq;
    const SlotParams s = sp[slot];
    const double pi = 3.1415926535897932;
    const int ix = (int)(m % g.Nx), iy = (int)(m / g.Nx);
    const int Nx = g.Nx, Ny = g.Ny;
    const double k0 = s.k0;
    const double kx = (ix < double(Nx / 2.)) ? 2 * pi * ix / (dx * Nx) : -2 * pi * (Nx - ix) / (dx * Nx);
    const double ky = (iy < double(Ny / 2.)) ? 2 * pi * iy / (dy * Ny) : -2 * pi * (Ny - iy) / (dy * Ny);
    const double qx0 = qxy[2 * iq], qy0 = qxy[2 * iq + 1];
    const double q0 = kx / k0 + qx0, q1 = ky / k0 + qy0;
    const double qn = sqrt(q0 * q0 + q1 * q1);
    const double qns = qn * qn;
    const double q0n = sqrt(qx0 * qx0 + qy0 * qy0);  // |q0|, as the reference (:252)
    c64 F[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    if (qn < NA) {
        c64 T[4][4];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) T[i][j] = mk(i == j ? 1.0 : 0.0, 0.0);
        for (int l = 0; l < n_iso; l++) {
            const double np = iso_index[l];
            const double nn = (l + 1 < n_iso) ? iso_index[l + 1] : s.nout_filter;
            const double neff = sqrt(np * np - qns), nneff = sqrt(nn * nn - qns);
            const double im_[2][2] = {{neff + q0 * q0 / neff, q0 * q1 / neff}, {q0 * q1 / neff, neff + q1 * q1 / neff}};
            const double ni[2][2] = {{(1 - (q0 / nn) * (q0 / nn)) / nneff, -q0 * q1 / (nn * nn * nneff)},
                                     {-q0 * q1 / (nn * nn * nneff), (1 - (q1 / nn) * (q1 / nn)) / nneff}};
            double tp[2][2], tm[2][2];
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    const double prod = ni[i][0] * im_[0][j] + ni[i][1] * im_[1][j];
                    tp[i][j] = 0.5 * ((i == j ? 1. : 0.) + prod);
                    tm[i][j] = (i == j ? 1. : 0.) - tp[i][j];
                }
            const double phase = (neff - sqrt(np * np - q0n)) * k0 * iso_thick[l];
            const c64 gph = mk(cos(phase), sin(phase)), gc = conj(gph);
            c64 Lm[4][4];
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    Lm[i][j] = tp[i][j] * gph;
                    Lm[i][j + 2] = tm[i][j] * gc;
                    Lm[i + 2][j] = tm[i][j] * gph;
                    Lm[i + 2][j + 2] = tp[i][j] * gc;
                }
            c64 Tn[4][4];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    c64 a = mk(0.0, 0.0);
                    for (int k = 0; k < 4; k++) a = cmad(a, Lm[i][k], T[k][j]);
                    Tn[i][j] = a;
                }
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) T[i][j] = Tn[i][j];
        }
        const c64 det = T[2][2] * T[3][3] - T[2][3] * T[3][2];
        const c64 id = crcp(det);
        const c64 inv[2][2] = {{T[3][3] * id, -(T[2][3] * id)}, {-(T[3][2] * id), T[2][2] * id}};
        const double nout = s.nout_filter;
        const double fz = (sqrt(nout * nout - qns) - sqrt(nout * nout - q0n)) * k0 * z_foc;
        const c64 ph = mk(cos(fz), sin(fz));
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                c64 a = T[i][j];
                for (int k = 0; k < 2; k++)
                    for (int l2 = 0; l2 < 2; l2++) a = a - T[i][k + 2] * inv[k][l2] * T[l2 + 2][j];
                F[2 * i + j] = a * ph;
            }
    }
    const double norm = double(Nx * (int64_t)Ny);
    for (int pol = 0; pol < 2; pol++) {
        c64 *e0 = E + (((int64_t)(slot * nq + iq) * 2 + pol) * 2) * g.Nfull + m;
        c64 *e1 = e0 + g.Nfull;
        const c64 x = *e0, y = *e1;
        const c64 a = F[0] * x + F[1] * y, b = F[2] * x + F[3] * y;
        *e0 = mk(a.x / norm, a.y / norm);
        *e1 = mk(b.x / norm, b.y / norm);
    }
}

void launch_filter(c64 *E, const Geo &g, const SlotParams *sp, int nslots, const double *qxy, int nq,
                   const double *iso_thick, const double *iso_index, int n_iso, double dx, double dy, double z_foc,
                   double NA, hipStream_t st) {
    hipLaunchKernelGGL(k_filter, dim3(blocks_for(g.Nfull, 256), nslots * nq), dim3(256), 0, st, E, g, sp, qxy, nq,
                       iso_thick, iso_index, n_iso, dx, dy, z_foc, NA);
    NMK_LAUNCH_CHECK();
}

// explicit instantiations
#define NMK_INST(T)                                                                                                   \
    template void launch_slab_coefs<T>(const double *, const uint8_t *, const double *, const Geo &,                \
                                       const SlotParams *, int, int, int, C2<T> *, int64_t, C2<T> *, int64_t, int, \
                                       const LinePlan &, hipStream_t);                                              \
    template void launch_init_fields<T>(const double *, const uint8_t *, const Geo &, const SlotParams *,             \
                                        const PropParams *, int, int, double, double, double, C2<T> *, hipStream_t);  \
    template void launch_pack_fields<T>(const c64 *, const Geo &, const PropParams *, int, int, C2<T> *, hipStream_t); \
    template void launch_boundary<T>(const C2<T> *, C2<T> *, const C2<T> *, const C2<T> *, const Geo &,             \
                                     const PropParams *, int, int, bool, hipStream_t);                                      \
    template void launch_explicit<T>(int, const C2<T> *, C2<T> *, const C2<T> *, const Geo &, const PropParams *,    \
                                     int, hipStream_t);                                                               \
    template void launch_bulk_plane<T>(const C2<T> *, const Geo &, const SlotParams *, const PropParams *, int, \
                                       int, int, c64 *, hipStream_t);                                           \
    template void launch_xsolve<T>(int, const C2<T> *, C2<T> *, const C2<T> *, C2<T> *, const C2<T> *,              \
                                   const C2<T> *, const Geo &, const PropParams *, int, const LinePlan &, int, int,  \
                                   int, int, hipStream_t);                                                            \
    template void launch_ysolve<T>(int, const C2<T> *, C2<T> *, const C2<T> *, C2<T> *, const C2<T> *, const Geo &, \
                                   const PropParams *, int, const LinePlan &, int, int, int, int, hipStream_t);       \
    template size_t xsolve_lds_bytes<T>(const LinePlan &, int, int);                                                  \
    template size_t ysolve_lds_bytes<T>(const LinePlan &, int, int);                                                  \
    template void launch_output<T>(const C2<T> *, const double *, const uint8_t *, const Geo &, const SlotParams *,  \
                                   const PropParams *, int, int, c64 *, hipStream_t);
NMK_INST(float)
NMK_INST(double)

}  // namespace nmk
