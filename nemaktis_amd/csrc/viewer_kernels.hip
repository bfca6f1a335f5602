// HIP kernels of the microscope imaging step (objective refocus and
// micrograph synthesis) for CDNA4 (gfx950).  Every kernel is bound by HBM
// bandwidth: consecutive threads walk consecutive x, the refocus reads and
// writes each field value three times (rows, columns, rows), the spectrum
// pass reads each focused value once.
#include <cmath>
#include <stdexcept>
#include <string>

#include "fft_device.h"
#include "kernels.h"
#include "viewer.h"

namespace nmk {

#define NMK_VIEW_LAUNCH_CHECK()                                                                  \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP launch failed: ") +     \
                                                       hipGetErrorString(e_));                   \
    } while (0)

namespace {
constexpr double kPi = 3.141592653589793;
inline unsigned blocks_of(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

// ---------------------------------------------------------------------------
// objective filter of one (wavelength, wavevector) plane, with the 1/N of the
// backward transform folded in (light_propagator.py:606-623, 710-719):
//   k = (kx + k0 qx, ky + k0 qy), kn = 2 pi n_out / lambda
//   mask   = |k|^2 < (k0 NA)^2
//   filter = exp(i (sqrt(kn^2 - |k|^2) - kn))  for |k|^2 < kn^2, else 0
//   F      = mask * filter^|zf|, conjugated for zf < 0
// filter^|zf| is numpy's principal power: exp(i |zf| arg(filter)), arg in
// (-pi, pi], and 0^0 = 1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_view_filter(c64 *__restrict__ filt, int Nx, int Ny, double dx, double dy,
                                                     const ViewPlane *__restrict__ planes, double NA, double zf,
                                                     double norm) {
    const int64_t N = (int64_t)Nx * Ny;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    const ViewPlane p = planes[blockIdx.y];
    const int ix = (int)(m % Nx), iy = (int)(m / Nx);
    const double kx = (fmod(ix + 0.5 * Nx, (double)Nx) - 0.5 * Nx) * 2 * kPi / (Nx * dx);
    const double ky = (fmod(iy + 0.5 * Ny, (double)Ny) - 0.5 * Ny) * 2 * kPi / (Ny * dy);
    const double ux = kx + p.px, uy = ky + p.py;
    const double kSqr = ux * ux + uy * uy;
    const double kNA = p.k0 * NA;
    c64 v = mk(0.0, 0.0);
    if (kSqr < kNA * kNA) {
        if (zf == 0) {
            v = mk(norm, 0.0);
        } else if (kSqr < p.kn * p.kn) {
            const double th = sqrt(p.kn * p.kn - kSqr) - p.kn;
            double s, c;
            sincos(th, &s, &c);
            const double ph = fabs(zf) * atan2(s, c);
            sincos(ph, &s, &c);
            v = mk(norm * c, zf < 0 ? -norm * s : norm * s);
        }
    }
    filt[(int64_t)blockIdx.y * N + m] = v;
}

void launch_view_filter(c64 *filt, int Nx, int Ny, double dx, double dy, const ViewPlane *planes, int nplanes,
                        double NA, double zf, hipStream_t st) {
    const int64_t N = (int64_t)Nx * Ny;
    hipLaunchKernelGGL(k_view_filter, dim3(blocks_of(N, 256), nplanes), dim3(256), 0, st, filt, Nx, Ny, dx, dy,
                       planes, NA, zf, 1.0 / (double)N);
    NMK_VIEW_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// column pass of the refocus, in place: forward transform of C adjacent
// columns, product with the filter plane of their (wavelength, wavevector),
// backward transform.  Work items run component-fastest and each XCD takes a
// contiguous range, so the four components reading one filter tile share L2.
// ---------------------------------------------------------------------------
template <int C, bool ODD, int NT = 256>
__global__ __launch_bounds__(NT) void k_view_columns(c64 *__restrict__ data, const c64 *__restrict__ filt, int Nx,
                                                      int ngroups, int64_t nitems, FftPlan pl,
                                                      const c64 *__restrict__ tw) {
    extern __shared__ c64 fbuf[];
    const int64_t item = xcd_tile(blockIdx.x, gridDim.x);
    if (item >= nitems) return;
    const int n = pl.n;
    const int comp = (int)(item & 3);
    const int64_t rest = item >> 2;
    const int g = (int)(rest % ngroups);
    const int64_t lq = rest / ngroups;
    const int64_t N = (int64_t)Nx * n;
    c64 *plane = data + (lq * 4 + comp) * N;
    const c64 *fp = filt + lq * N;
    const int x0 = g * C;
    c64 *stw = fbuf + 2 * C * n;
    // the filter values this thread multiplies, loaded with the fields (n * C <= 2048)
    constexpr int kPer = 2048 / NT;
    c64 fr[kPer];
#pragma unroll
    for (int e = 0; e < kPer; e++) {
        const int t = threadIdx.x + e * NT;
        const int i = t % C, y = t / C;
        fr[e] = (t < n * C && x0 + i < Nx) ? fp[(int64_t)y * Nx + x0 + i] : mk(0.0, 0.0);
    }
    for (int t = threadIdx.x; t < n * C; t += blockDim.x) {
        const int i = t % C, y = t / C;
        fbuf[t] = (x0 + i < Nx) ? plane[(int64_t)y * Nx + x0 + i] : mk(0.0, 0.0);
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
    __syncthreads();
    c64 *a = fft_stages<C, false, ODD>(fbuf, fbuf + n * C, pl, stw, -1);
#pragma unroll
    for (int e = 0; e < kPer; e++) {
        const int t = threadIdx.x + e * NT;
        if (t < n * C) a[t] = a[t] * fr[e];
    }
    __syncthreads();
    const c64 *r = fft_stages<C, false, ODD>(a, a == fbuf ? fbuf + n * C : fbuf, pl, stw, +1);
    for (int t = threadIdx.x; t < n * C; t += blockDim.x) {
        const int i = t % C, y = t / C;
        if (x0 + i < Nx) plane[(int64_t)y * Nx + x0 + i] = r[t];
    }
}

// fields of the 4 components of every (wavelength, wavevector) times its
// filter plane (the column pass of lengths without an LDS plan)
__global__ __launch_bounds__(256) void k_view_apply_filter(c64 *__restrict__ data, const c64 *__restrict__ filt,
                                                           int64_t N) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    const int64_t plane = blockIdx.y;  // (wavelength, wavevector) * 4 + component
    data[plane * N + m] = data[plane * N + m] * filt[(plane >> 2) * N + m];
}

void launch_view_columns(c64 *data, const c64 *filt, int Nx, int Ny, int nlq, const c64 *tw_y, const FftPlan &py,
                         hipStream_t st) {
    if (py.direct) {
        // direct DFT of the columns, the filter, and the inverse DFT
        const int64_t N = (int64_t)Nx * Ny;
        launch_dft_lines(data, data, Ny, Nx, nlq * 4, N, 1, Nx, -1, tw_y, st);
        hipLaunchKernelGGL(k_view_apply_filter, dim3(blocks_of(N, 256), nlq * 4), dim3(256), 0, st, data, filt, N);
        NMK_VIEW_LAUNCH_CHECK();
        launch_dft_lines(data, data, Ny, Nx, nlq * 4, N, 1, Nx, +1, tw_y, st);
        return;
    }
    // power-of-two columns of at most 256 points: eight adjacent columns per
    // block of 512 threads, so that a row of the block is a whole 128-byte line
    const bool wide = fft_pow2(py) && 8 * Ny <= 2048 && Nx >= 8;
    const int C = wide ? 8 : fft_cols_per_block(Ny);
    if (Ny * C > 8 * 256) throw std::runtime_error("refocus columns: more than 2048 values per block");
    const int ngroups = (Nx + C - 1) / C;
    const int64_t nitems = (int64_t)ngroups * nlq * 4;
    const unsigned grid = (unsigned)((nitems + 7) / 8 * 8);
    const size_t lds = fft_lds_bytes(Ny, C);
    if (wide) {
        allow_kernel_lds(reinterpret_cast<const void *>(k_view_columns<8, false, 512>), lds);
        hipLaunchKernelGGL((k_view_columns<8, false, 512>), dim3(grid), dim3(512), lds, st, data, filt, Nx, ngroups,
                           nitems, py, tw_y);
        NMK_VIEW_LAUNCH_CHECK();
        return;
    }
    switch (C) {
#define NMK_VCOLS(CC)                                                                                         \
    case CC:                                                                                                  \
        if (fft_pow2(py)) {                                                                                   \
            allow_kernel_lds(reinterpret_cast<const void *>(k_view_columns<CC, false>), lds);                \
            hipLaunchKernelGGL((k_view_columns<CC, false>), dim3(grid), dim3(256), lds, st, data, filt, Nx,   \
                               ngroups, nitems, py, tw_y);                                                    \
        } else {                                                                                              \
            allow_kernel_lds(reinterpret_cast<const void *>(k_view_columns<CC, true>), lds);                 \
            hipLaunchKernelGGL((k_view_columns<CC, true>), dim3(grid), dim3(256), lds, st, data, filt, Nx,    \
                               ngroups, nitems, py, tw_y);                                                    \
        }                                                                                                     \
        break;
        NMK_VCOLS(8)
        NMK_VCOLS(4)
        NMK_VCOLS(2)
        NMK_VCOLS(1)
#undef NMK_VCOLS
    default: throw std::runtime_error("refocus columns: bad block shape");
    }
    NMK_VIEW_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// q-averaged spectrum of one wavelength (field_viewer.py:353-446):
//   S = [[Ex(x), Ex(y)], [Ey(x), Ey(y)]]   (sample_transfer_matrix)
//   T = A (Wu (S R)),  s(q) = |T|_F^2 / 2,  spec = sum_q w_q s(q)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mat2(const c64 *a, const c64 *b, c64 *o) {
    o[0] = a[0] * b[0] + a[1] * b[2];
    o[1] = a[0] * b[1] + a[1] * b[3];
    o[2] = a[2] * b[0] + a[3] * b[2];
    o[3] = a[2] * b[1] + a[3] * b[3];
}

__global__ __launch_bounds__(256) void k_view_spectrum(const c64 *__restrict__ E, int64_t N, int Nq, int q0, int q1,
                                                       const double *__restrict__ w,
                                                       const JonesSet *__restrict__ jones,
                                                       double *__restrict__ spec) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    const int l = blockIdx.y;
    const JonesSet J = jones[l];
    double acc = 0;
#pragma unroll 4
    for (int q = q0; q < q1; q++) {
        const c64 *e = E + ((int64_t)l * Nq + q) * 4 * N + m;
        const c64 S[4] = {e[0], e[2 * N], e[N], e[3 * N]};
        c64 t1[4], t2[4], T[4];
        mat2(S, J.R, t1);
        mat2(J.Wu, t1, t2);
        mat2(J.A, t2, T);
        double s = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) s += T[k].x * T[k].x + T[k].y * T[k].y;
        acc += 0.5 * s * w[q];
    }
    spec[(int64_t)l * N + m] = acc;
}

void launch_view_spectrum(const c64 *E, int64_t N, int Nl, int Nq, int q0, int q1, const double *w,
                          const JonesSet *jones, double *spec, hipStream_t st) {
    hipLaunchKernelGGL(k_view_spectrum, dim3(blocks_of(N, 256), Nl), dim3(256), 0, st, E, N, Nq, q0, q1, w, jones,
                       spec);
    NMK_VIEW_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// colour of one pixel: XYZ = sum_l intensity spec_l cmf_l, then linear sRGB
// (D65 primaries) or the Y channel for grayscale, then the sRGB transfer
// function with clipping (gamma) or a clip to [0, 1]; also writes the
// spectrum in the (Ny, Nx, Nl) layout of FieldViewer.get_spectrum.
// ---------------------------------------------------------------------------
__host__ __device__ inline double srgb_gamma_fn(double v) {
    if (v > 0.0031308) return v > 1. ? 1. : 1.055 * pow(v, 1. / 2.4) - 0.055;
    return v < 0. ? 0. : 12.92 * v;
}
double srgb_gamma(double v) { return srgb_gamma_fn(v); }

__global__ __launch_bounds__(256) void k_view_colour(const double *__restrict__ spec, int64_t N, int Nl,
                                                     const double *__restrict__ cmf, double intensity, int grayscale,
                                                     int gamma, double *__restrict__ spec_out,
                                                     double *__restrict__ image) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    double X = 0, Y = 0, Z = 0;
    for (int l = 0; l < Nl; l++) {
        const double s = spec[(int64_t)l * N + m];
        spec_out[m * Nl + l] = s;
        if (image) {
            const double v = intensity * s;
            X += v * cmf[3 * l];
            Y += v * cmf[3 * l + 1];
            Z += v * cmf[3 * l + 2];
        }
    }
    if (!image) return;
    double rgb[3];
    if (grayscale) {
        rgb[0] = rgb[1] = rgb[2] = Y;
    } else {
        rgb[0] = 3.2404542 * X - 1.5371385 * Y - 0.4985314 * Z;
        rgb[1] = -0.9692660 * X + 1.8760108 * Y + 0.0415560 * Z;
        rgb[2] = 0.0556434 * X - 0.2040259 * Y + 1.0572252 * Z;
    }
    for (int c = 0; c < 3; c++)
        image[m * 3 + c] = gamma ? srgb_gamma_fn(rgb[c]) : fmin(1., fmax(0., rgb[c]));
}

void launch_view_colour(const double *spec, int64_t N, int Nl, const double *cmf, double intensity, int grayscale,
                        int gamma, double *spec_out, double *image, hipStream_t st) {
    hipLaunchKernelGGL(k_view_colour, dim3(blocks_of(N, 256)), dim3(256), 0, st, spec, N, Nl, cmf, intensity,
                       grayscale, gamma, spec_out, image);
    NMK_VIEW_LAUNCH_CHECK();
}

}  // namespace nmk
