// Complex arithmetic and the parameter blocks shared by the HIP kernels of
// the bpm-solver march.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nmk {

template <class T>
struct alignas(2 * sizeof(T)) C2 {
    T x, y;
};
using c32 = C2<float>;
using c64 = C2<double>;

template <class T>
__host__ __device__ __forceinline__ C2<T> mk(T a, T b) {
    return C2<T>{a, b};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> operator+(C2<T> a, C2<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> operator-(C2<T> a, C2<T> b) {
    return {a.x - b.x, a.y - b.y};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> operator-(C2<T> a) {
    return {-a.x, -a.y};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> operator*(C2<T> a, C2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> operator*(T s, C2<T> a) {
    return {s * a.x, s * a.y};
}
// acc + a*b
template <class T>
__host__ __device__ __forceinline__ C2<T> cmad(C2<T> acc, C2<T> a, C2<T> b) {
    return {acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> conj(C2<T> a) {
    return {a.x, -a.y};
}
// 1/a = conj(a)/|a|^2 with one real reciprocal (the line-solver pivots are
// O(1), far from the overflow range); FP32 uses the hardware reciprocal
// (1 ulp), FP64 an IEEE division
__host__ __device__ __forceinline__ float real_rcp(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(v);
#else
    return 1.0f / v;
#endif
}
__host__ __device__ __forceinline__ double real_rcp(double v) { return 1.0 / v; }
template <class T>
__host__ __device__ __forceinline__ C2<T> crcp(C2<T> a) {
    const T r = real_rcp(a.x * a.x + a.y * a.y);
    return {a.x * r, -a.y * r};
}
template <class T>
__host__ __device__ __forceinline__ C2<T> cdiv(C2<T> a, C2<T> b) {
    return a * crcp(b);
}
template <class T, class U>
__host__ __device__ __forceinline__ C2<T> cast2(C2<U> a) {
    return {(T)a.x, (T)a.y};
}

// Under round-robin dispatch XCD k receives blocks k, k+8, k+16, ...; this
// remap gives them consecutive work items (nbx = gridDim.x, a multiple of 8),
// so items that read the same data (neighbouring rows, the components of one
// plane) run on one XCD and share its L2.
__device__ __forceinline__ int xcd_tile(int bx, int nbx) { return (bx & 7) * (nbx >> 3) + (bx >> 3); }

// ---- launch-time parameter blocks ----------------------------------------

// geometry of the run: full grid (with the duplicate column/row) and the
// march grid (periodic interior along axes with more than 3 points)
struct Geo {
    int Nx, Ny, Nz;
    int Mx, My;       // march grid
    int px, py;       // 1 if the axis is periodic (N > 3)
    int field_dim;    // 3 director, 6 Q-tensor
    int64_t Nfull;    // Nx*Ny
    int64_t NP;       // Mx*My
};

// per-wavelength scalars of one batch slot
struct SlotParams {
    double k0, dX, dY, dZ;
    double no, ne, ne_imag, nhost;
    double eperp, ea, ehost;
    double alpha_im;      // absorption factor alpha = 1 + i*alpha_im
    double nin, nout_fresnel, nout_filter;
    double nref;          // filled on the device by the reduction
    int iwl;              // wavelength index in the caller's array
    int pad;
};

// per-propagation (wavelength slot, wavevector, polarisation) scalars
struct PropParams {
    int slot, iq, pol, pad;
    int ox, oy;           // first tap offset of the Catmull-Rom shift (floor(s)-1)
    double wx[4], wy[4];  // Catmull-Rom weights along x and y
    int o1x, o1y, o2x, o2y;      // first tap of one shift / of two composed shifts
    double w1x[7], w1y[7];       // one shift as a 7-tap filter (last 3 taps zero)
    double w2x[7], w2y[7];       // two shifts composed: w * w
    double qx, qy;
    double ex, ey;        // input polarisation (cos, sin of pol*pi/2)
    double ph_re, ph_im;  // accumulated global phase of all shift half-steps
    double ph_half;       // phase angle of one shift half-step, -dZ (qx^2 + qy^2) / (4 nref)
};

// chunking of one periodic line of M unknowns into P chunks (P a power of
// two) of lengths base+1 (the first rem chunks) or base; lmax bounds the
// chunk length and sets the register arrays of the chunk solver
struct LinePlan {
    int M, P, base, rem, lmax;
};
__host__ __device__ __forceinline__ int chunk_start(const LinePlan &lp, int p) {
    return p * lp.base + (p < lp.rem ? p : lp.rem);
}
__host__ __device__ __forceinline__ int chunk_len(const LinePlan &lp, int p) { return lp.base + (p < lp.rem ? 1 : 0); }
// chunk p and offset i of line element x
__host__ __device__ __forceinline__ void chunk_of(const LinePlan &lp, int x, int &p, int &i) {
    const int big = lp.rem * (lp.base + 1);
    if (x < big) {
        p = x / (lp.base + 1);
        i = x - p * (lp.base + 1);
    } else {
        p = lp.rem + (x - big) / lp.base;
        i = x - big - (p - lp.rem) * lp.base;
    }
}

// coefficient plane slots (per interior point, stored as mu*R and P)
enum : int {
    CF_DX = 0,     // 4 blocks x {minus, centre, plus}
    CF_DY = 12,    // 4 blocks x {minus, centre, plus}
    CF_XY = 24,    // 4 blocks x {mxmy, pxmy, mxpy, pxpy}
    CF_NMU = 40    // mu*R planes
};
// mu*R = i (dZ/2) R.  R is real apart from the walk-off entries, which sit
// in DX blocks 00, 01, 10 and DY blocks 01, 10, 11 at the minus and plus
// slots (W_ab = (v_a d_b + d_a v_b) / 2 has no centre term and no DX part in
// block 11 or DY part in block 00): those 12 planes are stored complex, and
// so are DY00 and DY11 whole (the y-line rows, which a y-line block reads for
// a component it knows only at run time); the other 24 planes hold the
// imaginary part of mu*R only (their real part is zero).
__host__ __device__ constexpr bool cf_complex(int k) {
    const int fam = k < CF_DY ? 0 : 1, blk = (k - (fam ? CF_DY : CF_DX)) / 3, sl = k % 3;
    return k < CF_XY && ((sl != 1 && blk != (fam ? 0 : 3)) || (fam == 1 && (blk == 0 || blk == 3)));
}
// index of plane k inside its region (complex or imaginary-only)
__host__ __device__ constexpr int cf_index(int k) {
    int n = 0;
    for (int j = 0; j < k; j++) n += cf_complex(j) == cf_complex(k);
    return n;
}
constexpr int CF_NCPLX = 16 + 3;  // complex planes: the 16 complex mu*R planes, then P
constexpr int CF_P = 16;          // complex plane of op_xx (op_yy, op_xy follow)
constexpr int CF_NIMAG = 24;      // imaginary-only planes
// Coefficient planes are stored in tiles of kCfTw x kCfTh = 16 x 4 march-grid
// points (row-major in the tile): per tile the CF_NCPLX complex planes, then
// the CF_NIMAG imaginary-only planes (64 scalars each), so a wave reads every
// tap of its points as contiguous segments at offsets fixed at compile time.
// Tiles are row-major over the grid, ceil(Mx / kCfTw) per tile row.  A y-line
// wave (8 columns x 8 rows) then touches 2 tiles instead of 8 (64 x 1
// tiles), while x-line waves (64 columns of one row) still read whole
// 128-byte segments of every complex plane.
constexpr int kCfTile = 64;
constexpr int kCfTw = 16, kCfTh = kCfTile / kCfTw;
static_assert(kCfTw * kCfTh == kCfTile && (kCfTw & (kCfTw - 1)) == 0, "tile shape");
constexpr int kCfTileC2 = kCfTile * CF_NCPLX + kCfTile * CF_NIMAG / 2;  // one tile in complex elements
__host__ __device__ __forceinline__ int64_t cf_slot_size(int64_t Mx, int64_t My) {
    return ((Mx + kCfTw - 1) / kCfTw) * ((My + kCfTh - 1) / kCfTh) * kCfTileC2;
}
// position of point (x, y) inside its tile
__host__ __device__ __forceinline__ unsigned cf_lane(unsigned x, unsigned y) {
    return (y % kCfTh) * kCfTw + x % kCfTw;
}
// offset of complex plane 0 of point (x, y) inside its slot (slots stay below
// 2^32 elements: at most 8.6e6 tiles)
__host__ __device__ __forceinline__ unsigned cf_point(unsigned x, unsigned y, unsigned Mx) {
    const unsigned tpr = (Mx + kCfTw - 1) / kCfTw;
    return ((y / kCfTh) * tpr + x / kCfTw) * (unsigned)kCfTileC2 + cf_lane(x, y);
}
// mu*R plane K at the point whose complex plane 0 is at pc (lane = cf_lane of the point)
template <int K, class T>
__device__ __forceinline__ C2<T> cf_load(const C2<T> *pc, unsigned lane) {
    constexpr int idx = cf_index(K);  // evaluated at compile time
    if constexpr (cf_complex(K)) {
        return pc[idx * kCfTile];
    } else {
        const T *pi = reinterpret_cast<const T *>(pc + CF_NCPLX * kCfTile) - lane;
        return C2<T>{(T)0, pi[idx * kCfTile]};
    }
}

// slab-material fields the coefficients read (FP64, per point of a tile)
enum : int {
    SM_EINV = 0, SM_KXX, SM_KYY, SM_KXY, SM_IXX, SM_IYY, SM_IXY, SM_VX, SM_VY,
    SM_COUNT
};

}  // namespace nmk
