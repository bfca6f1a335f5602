// Device side of the mixed-radix FFT used by the screen filter and the
// objective refocus: Stockham autosort stages over C lines held together in
// LDS, ping-ponging between two buffers.  Element k of line i sits at
// k*C + i, so the C lines of a block are read from HBM as rows of C
// consecutive complex values (columns of a plane) and the threads of a wave
// touch consecutive LDS words.  Stage with radix r after sub-transforms of
// length L:
//   y'[j + L p + L r k] = sum_q w_r^{pq} ( w_{Lr}^{jq} y[j + L k + (n/r) q] )
// Twiddles tw[t] = exp(-2 pi i t / n); sign < 0 is the forward transform
// (exp(-i)), sign > 0 the unnormalised backward one.
#pragma once

#include "device_common.h"
#include "kernels.h"

namespace nmk {

// one stage of radix R over the C lines of the block.  LM selects the LDS
// layout: line-major (element k of line i at i*n + k, rows) or interleaved
// (k*C + i, column tiles); work items run along the contiguous direction.
template <int C, bool LM, int R>
__device__ __forceinline__ void fft_stage(const c64 *__restrict__ b0, c64 *__restrict__ b1, int n, int L,
                                          const c64 *__restrict__ tw, int sign) {
    const int Lr = L * R, nb = n / R, tstep = n / Lr, rstep = n / R;
    auto W = [&](int idx) -> c64 {
        const c64 w = tw[idx];
        return sign > 0 ? conj(w) : w;
    };
    // R-th roots of unity of the butterfly
    c64 wr[R];
#pragma unroll
    for (int m = 0; m < R; m++) wr[m] = (R == 2 || R == 4) ? mk(1.0, 0.0) : W(m * rstep);
    const int step = LM ? 1 : C;          // LDS distance of consecutive elements of a line
    for (int t = threadIdx.x; t < nb * C; t += blockDim.x) {
        const int i = LM ? t / nb : t % C;
        const int b = LM ? t - i * nb : t / C;
        const int base = LM ? i * n : i;
        const int j = b % L, k = b / L;
        c64 u[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            const c64 v = b0[base + (j + L * k + rstep * q) * step];
            u[q] = (q == 0 || j == 0) ? v : v * W(j * q * tstep);
        }
        c64 *o = b1 + base + (j + Lr * k) * step;
        const int os = L * step;
        if constexpr (R == 2) {
            o[0] = u[0] + u[1];
            o[os] = u[0] - u[1];
        } else if constexpr (R == 4) {
            // w_4 = -i (forward) or +i (backward)
            const c64 t0 = u[0] + u[2], t1 = u[0] - u[2], t2 = u[1] + u[3], d = u[1] - u[3];
            const c64 t3 = sign > 0 ? mk(-d.y, d.x) : mk(d.y, -d.x);
            o[0] = t0 + t2;
            o[os] = t1 + t3;
            o[2 * os] = t0 - t2;
            o[3 * os] = t1 - t3;
        } else if constexpr (R == 8) {
            // two 4-point transforms of the even and odd inputs, joined by
            // w_8^k (k = 0..3; w_8^2 = -+i, w_8^1 and w_8^3 from the table)
            auto dft4 = [&](c64 a, c64 b, c64 c, c64 d, c64 (&y)[4]) {
                const c64 t0 = a + c, t1 = a - c, t2 = b + d, e = b - d;
                const c64 t3 = sign > 0 ? mk(-e.y, e.x) : mk(e.y, -e.x);
                y[0] = t0 + t2;
                y[1] = t1 + t3;
                y[2] = t0 - t2;
                y[3] = t1 - t3;
            };
            c64 ev[4], od[4];
            dft4(u[0], u[2], u[4], u[6], ev);
            dft4(u[1], u[3], u[5], u[7], od);
            const c64 o1 = od[1] * wr[1], o3 = od[3] * wr[3];
            const c64 o2 = sign > 0 ? mk(-od[2].y, od[2].x) : mk(od[2].y, -od[2].x);
            o[0] = ev[0] + od[0];
            o[os] = ev[1] + o1;
            o[2 * os] = ev[2] + o2;
            o[3 * os] = ev[3] + o3;
            o[4 * os] = ev[0] - od[0];
            o[5 * os] = ev[1] - o1;
            o[6 * os] = ev[2] - o2;
            o[7 * os] = ev[3] - o3;
        } else {
#pragma unroll
            for (int pq = 0; pq < R; pq++) {
                c64 a = u[0];
#pragma unroll
                for (int q = 1; q < R; q++) a = cmad(a, u[q], wr[(pq * q) % R]);
                o[pq * os] = a;
            }
        }
    }
}

// runs every stage of plan pl on the C lines in b0 (all threads of the block
// take part; the caller has synchronised after filling b0 and the twiddle
// table tw, normally an LDS copy) and returns the buffer that holds the result
// (ODD: the plan may hold radices 3 to 13; without it only 2, 4 and 8 are
// compiled, which keeps the power-of-two kernels small)
template <int C, bool LM, bool ODD>
__device__ c64 *fft_stages(c64 *b0, c64 *b1, const FftPlan &pl, const c64 *__restrict__ tw, int sign) {
    const int n = pl.n;
    int L = 1;
    for (int st = 0; st < pl.nst; st++) {
        const int r = (int)((pl.radices >> (4 * st)) & 15);
        if (r == 4) fft_stage<C, LM, 4>(b0, b1, n, L, tw, sign);
        else if (r == 8) fft_stage<C, LM, 8>(b0, b1, n, L, tw, sign);
        else if (r == 2) fft_stage<C, LM, 2>(b0, b1, n, L, tw, sign);
        else if constexpr (ODD) {
            switch (r) {
            case 3: fft_stage<C, LM, 3>(b0, b1, n, L, tw, sign); break;
            case 5: fft_stage<C, LM, 5>(b0, b1, n, L, tw, sign); break;
            case 7: fft_stage<C, LM, 7>(b0, b1, n, L, tw, sign); break;
            case 11: fft_stage<C, LM, 11>(b0, b1, n, L, tw, sign); break;
            default: fft_stage<C, LM, 13>(b0, b1, n, L, tw, sign); break;
            }
        }
        __syncthreads();
        c64 *tmp = b0;
        b0 = b1;
        b1 = tmp;
        L *= r;
    }
    return b0;
}

// true when every stage of the plan has radix 2, 4 or 8
__host__ __device__ inline bool fft_pow2(const FftPlan &pl) {
    for (int st = 0; st < pl.nst; st++) {
        const int r = (int)((pl.radices >> (4 * st)) & 15);
        if (r != 2 && r != 4 && r != 8) return false;
    }
    return true;
}

// lines per block: at least 1024 values per block of 256 threads (one
// radix-4 butterfly per thread and stage)
__host__ __device__ constexpr int fft_lines_per_block(int n, int cap) {
    int c = 1;
    while (2 * c <= cap && 2 * c * n <= 1024) c *= 2;
    return c;
}
// LDS bytes of a block: two line buffers and the twiddle table
__host__ __device__ constexpr size_t fft_lds_bytes(int n, int C) { return (size_t)(2 * C + 1) * n * 16; }

}  // namespace nmk
