/*
 * TEST INFRASTRUCTURE - NOT PRODUCT CODE.  FP64 CPU restatement of the
 * Nemaktis bpm-solver propagation step; see bpm_oracle.h for who may use it.
 *
 * Reference files followed (paths relative to the reference repository,
 * BPL = BeamPropagationSolver/src/BeamPropagationLib, PHY = .../PhysicsLib,
 * PP = .../PostprocessorLib):
 *   BPL/BPMIteration.cpp:53-177           loop order, field init, Fresnel placement
 *   BPL/BaseBPMOperator.cpp:3-13          k0, deltaX/Y/Z
 *   PHY/PermittivityTensorField.cpp:3-99  eps, K, sqrt(K), nref
 *   BPL/PhaseEvolutionOperator.cpp:3-63   phase operator
 *   BPL/SimpleShiftOperator.cpp:3-93      Catmull-Rom shift operator
 *   BPL/ParaxialPrimaryEvolutionOperator.cpp:4-490  D0, D1, D, R, ADI apply
 *   BPL/BlockSparseMatrix.cpp:18-550      sparsity, products, cyclic solves
 *   BPL/FresnelOperator.cpp:16-69         input/output Fresnel
 *   PP/ScreenOutput.cpp:13-39,184-302     screen filter
 */
#include "bpm_oracle.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

typedef std::complex<double> cd;
const double kPi = 3.1415926535897932;  // PHY/PhysicsCoefficients.h:14

enum { TX = 0, TY = 1, TXY = 2 };

// ---------------------------------------------------------------------------
// FFT used by the screen filter: iterative radix-2 Cooley-Tukey for powers of
// two, Bluestein's chirp-z algorithm (through a radix-2 FFT) otherwise.
// Unnormalised; sign -1 = forward, +1 = backward (FFTW convention).
// ---------------------------------------------------------------------------
void fft_pow2(std::vector<cd> &a, int sign) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2;
        for (size_t k = 0; k < half; k++) {
            double ang = sign * 2.0 * kPi * double(k) / double(len);
            cd w(std::cos(ang), std::sin(ang));
            for (size_t i = 0; i < n; i += len) {
                cd u = a[i + k], v = a[i + k + half] * w;
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
        }
    }
}

void fft_any(std::vector<cd> &a, int sign) {
    const size_t n = a.size();
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) { fft_pow2(a, sign); return; }
    size_t m = 1;
    while (m < 2 * n - 1) m <<= 1;
    std::vector<cd> chirp(n), A(m, cd(0, 0)), B(m, cd(0, 0));
    for (size_t k = 0; k < n; k++) {
        // exp(sign * i * pi * k^2 / n), with k^2 reduced mod 2n for accuracy
        unsigned long long k2 = (unsigned long long)k * k % (2ULL * n);
        double ang = sign * kPi * double(k2) / double(n);
        chirp[k] = cd(std::cos(ang), std::sin(ang));
    }
    for (size_t k = 0; k < n; k++) A[k] = a[k] * chirp[k];
    B[0] = std::conj(chirp[0]);
    for (size_t k = 1; k < n; k++) B[k] = B[m - k] = std::conj(chirp[k]);
    fft_pow2(A, -1);
    fft_pow2(B, -1);
    for (size_t k = 0; k < m; k++) A[k] *= B[k];
    fft_pow2(A, +1);
    for (size_t k = 0; k < n; k++) a[k] = A[k] * chirp[k] / double(m);
}

void fft2_planes(int Nx, int Ny, int howmany, int sign, cd *data) {
    for (int h = 0; h < howmany; h++) {
        cd *plane = data + (size_t)h * Nx * Ny;
#pragma omp parallel for
        for (int iy = 0; iy < Ny; iy++) {
            std::vector<cd> line(plane + (size_t)iy * Nx, plane + (size_t)(iy + 1) * Nx);
            fft_any(line, sign);
            std::copy(line.begin(), line.end(), plane + (size_t)iy * Nx);
        }
#pragma omp parallel for
        for (int ix = 0; ix < Nx; ix++) {
            std::vector<cd> line(Ny);
            for (int iy = 0; iy < Ny; iy++) line[iy] = plane[ix + (size_t)Nx * iy];
            fft_any(line, sign);
            for (int iy = 0; iy < Ny; iy++) plane[ix + (size_t)Nx * iy] = line[iy];
        }
    }
}

// ---------------------------------------------------------------------------
// Oracle state for one run.
// ---------------------------------------------------------------------------
struct Oracle {
    OracleParams p;
    const double *lc;
    int Nx, Ny, Nz, N;

    // wavelength-dependent scalars
    int iwl = -1;
    double lambda = 0, k0 = 0, dX = 0, dY = 0, dZ = 0;
    cd mu, absorp;
    double no = 0, ne = 0, ne_imag = 0, nhost = 0, eperp = 0, ea = 0, ehost = 0;
    double nref = 0;
    int iz = 0;

    // cached eps planes (12 values per point)
    std::vector<double> plane[2];
    int plane_iz[2] = {-1, -1};

    // stencil sub-matrices: index ((block*ns + shift)*N + m)
    std::vector<double> D0[3], D1[3], D[3];
    std::vector<cd> R[3];
    std::vector<cd> opxx, opyy, opxy;
    std::vector<cd> tmp, tmp2;

    static int ns(int t) { return t == TXY ? 4 : 3; }
    size_t ci(int t, int b, int s, int m) const { return ((size_t)(b * ns(t) + s)) * N + m; }

    bool xper() const { return Nx > 3; }
    bool yper() const { return Ny > 3; }
    bool inrow(int t, int ix, int iy) const {
        if (t == TX) return xper() && ix < Nx - 1;
        if (t == TY) return yper() && iy < Ny - 1;
        return xper() && yper() && ix < Nx - 1 && iy < Ny - 1;
    }
    // stencil column along one axis (BPL/BlockSparseMatrix.cpp:86-88, 277-279, 470-473)
    int colx(int s, int ix) const {
        if (s == 0) return ix == 0 ? Nx - 2 : ix - 1;
        if (s == 1) return ix;
        return ix == Nx - 2 ? 0 : ix + 1;
    }
    int coly(int s, int iy) const {
        if (s == 0) return iy == 0 ? Ny - 2 : iy - 1;
        if (s == 1) return iy;
        return iy == Ny - 2 ? 0 : iy + 1;
    }
    // DXDY slots: 0 mxmy, 1 pxmy, 2 mxpy, 3 pxpy
    static int xy_sx(int s) { return (s & 1) ? 2 : 0; }
    static int xy_sy(int s) { return (s & 2) ? 2 : 0; }
    int col(int t, int s, int ix, int iy) const {
        if (t == TX) return colx(s, ix) + Nx * iy;
        if (t == TY) return ix + Nx * coly(s, iy);
        return colx(xy_sx(s), ix) + Nx * coly(xy_sy(s), iy);
    }

    Oracle(const OracleParams *pp, const double *lc_) : p(*pp), lc(lc_) {
        Nx = p.Nx; Ny = p.Ny; Nz = p.Nz; N = Nx * Ny;
        for (int t = 0; t < 3; t++) {
            D0[t].assign((size_t)4 * ns(t) * N, 0.0);
            D1[t].assign((size_t)4 * ns(t) * N, 0.0);
            D[t].assign((size_t)4 * ns(t) * N, 0.0);
            R[t].assign((size_t)4 * ns(t) * N, cd(0, 0));
        }
        opxx.resize(N); opyy.resize(N); opxy.resize(N);
        tmp.resize(2 * (size_t)N); tmp2.resize(2 * (size_t)N);
        plane[0].resize(12 * (size_t)N); plane[1].resize(12 * (size_t)N);
    }

    bool mask_at(size_t v) const { return p.mask ? (p.mask[v] >= 0) : true; }

    // PHY/PermittivityTensorField.cpp:18-40 (director), :41-65 (Q-tensor), :70-93
    double eps_voxel(size_t v, double *e) const {
        const int fd = p.field_dim;
        const double *n = lc + (size_t)fd * v;
        bool m = mask_at(v);
        if (fd == 3) {
            e[0] = m ? eperp + ea * (n[0] * n[0]) : ehost;
            e[1] = m ? eperp + ea * (n[1] * n[1]) : ehost;
            e[2] = m ? eperp + ea * (n[2] * n[2]) : ehost;
            e[3] = m ? ea * n[0] * n[1] : 0;
            e[4] = m ? ea * n[0] * n[2] : 0;
            e[5] = m ? ea * n[1] * n[2] : 0;
        } else {
            double ea_eff = 2 * ea / 3;
            double eiso = eperp + ea / 3.;
            e[0] = m ? eiso + ea_eff * n[0] : ehost;
            e[1] = m ? eiso + ea_eff * n[1] : ehost;
            e[2] = m ? eiso + ea_eff * n[2] : ehost;
            e[3] = m ? ea_eff * n[3] : 0;
            e[4] = m ? ea_eff * n[4] : 0;
            e[5] = m ? ea_eff * n[5] : 0;
        }
        e[6] = e[0] - (e[4] * e[4]) / e[2];
        e[7] = e[1] - (e[5] * e[5]) / e[2];
        e[8] = e[3] - e[4] * e[5] / e[2];
        double det = e[6] * e[7] - e[8] * e[8];
        double tr = 0.5 * (e[6] + e[7]);
        double neff;
        if (tr * tr - det < 1e-8)
            neff = std::sqrt(tr);
        else
            neff = (std::sqrt(tr + std::sqrt(tr * tr - det)) + std::sqrt(tr - std::sqrt(tr * tr - det))) / 2.;
        e[9] = neff + (e[6] - e[7]) / (4 * neff);
        e[10] = neff + (e[7] - e[6]) / (4 * neff);
        e[11] = e[8] / (2 * neff);
        return neff;
    }

    void compute_plane(int z, std::vector<double> &out) const {
#pragma omp parallel for
        for (int m = 0; m < N; m++) eps_voxel((size_t)m + (size_t)N * z, &out[12 * (size_t)m]);
    }
    // returns eps of plane z (12 per point), keeping at most two planes cached
    const double *eps_plane(int z) {
        for (int k = 0; k < 2; k++)
            if (plane_iz[k] == z) return plane[k].data();
        int victim = (plane_iz[0] == z - 1) ? 1 : 0;
        compute_plane(z, plane[victim]);
        plane_iz[victim] = z;
        return plane[victim].data();
    }

    void set_wavelength(int w) {
        iwl = w;
        lambda = p.wl[w];
        no = p.no[w]; ne = p.ne[w]; ne_imag = p.ne_imag[w]; nhost = p.nhost[w];
        eperp = no * no;
        ea = ne * ne - eperp;
        ehost = nhost * nhost;
        k0 = 2 * kPi / lambda;
        dX = k0 * p.dx; dY = k0 * p.dy; dZ = k0 * p.dz;
        mu = cd(0, dZ / 2.);
        absorp = cd(1, ne_imag / (ne - no));  // BPL/PhaseEvolutionOperator.cpp:12 (NaN if ne == no)
        plane_iz[0] = plane_iz[1] = -1;
        // nref = mean n_eff over every voxel (PHY/PermittivityTensorField.cpp:94-98),
        // summed plane by plane in a fixed order
        std::vector<double> part(Nz, 0.0);
#pragma omp parallel for
        for (int z = 0; z < Nz; z++) {
            double e[12], s = 0;
            for (int m = 0; m < N; m++) s += eps_voxel((size_t)m + (size_t)N * z, e);
            part[z] = s;
        }
        double sum = 0;
        for (int z = 0; z < Nz; z++) sum += part[z];
        nref = sum / ((double)Nx * Ny * Nz);
        build_D0();
        iz = 0;
    }

    // BPL/ParaxialPrimaryEvolutionOperator.cpp:70-99
    void build_D0() {
        for (int t = 0; t < 3; t++) std::fill(D0[t].begin(), D0[t].end(), 0.0);
        double iy2 = 1. / (dY * dY), ix2 = 1. / (dX * dX), ixy = 0.25 / (dX * dY);
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                int m = ix + Nx * iy;
                if (iy < Ny - 1 && Ny > 3) {
                    D0[TY][ci(TY, 0, 0, m)] = iy2;
                    D0[TY][ci(TY, 0, 1, m)] = -2. * iy2;
                    D0[TY][ci(TY, 0, 2, m)] = iy2;
                }
                if (iy < Ny - 1 && ix < Nx - 1 && Nx > 3 && Ny > 3) {
                    for (int b : {1, 2}) {
                        D0[TXY][ci(TXY, b, 3, m)] = -ixy;  // pxpy
                        D0[TXY][ci(TXY, b, 1, m)] = ixy;   // pxmy
                        D0[TXY][ci(TXY, b, 2, m)] = ixy;   // mxpy
                        D0[TXY][ci(TXY, b, 0, m)] = -ixy;  // mxmy
                    }
                }
                if (ix < Nx - 1 && Nx > 3) {
                    D0[TX][ci(TX, 3, 0, m)] = ix2;
                    D0[TX][ci(TX, 3, 1, m)] = -2. * ix2;
                    D0[TX][ci(TX, 3, 2, m)] = ix2;
                }
            }
    }

    // slab-averaged quantities of point m for slab iz
    struct Slab {
        double ezz_inv, kxx, kyy, kxy, sxx, syy, sxy, vx, vy;
    };
    Slab slab_at(const double *e0, const double *e1, int m) const {
        const double *a = e0 + 12 * (size_t)m, *b = e1 + 12 * (size_t)m;
        Slab s;
        s.ezz_inv = (1. / a[2] + 1. / b[2]) / 2;
        s.kxx = (a[6] + b[6]) / 2; s.kyy = (a[7] + b[7]) / 2; s.kxy = (a[8] + b[8]) / 2;
        s.sxx = (a[9] + b[9]) / 2; s.syy = (a[10] + b[10]) / 2; s.sxy = (a[11] + b[11]) / 2;
        s.vx = a[4] / a[2] + b[4] / b[2];  // summed, scaled later as in the reference
        s.vy = a[5] / a[2] + b[5] / b[2];
        return s;
    }

    // ParaxialPrimaryEvolutionOperator::update_D1 (:421-490)
    void update_D1(const double *e0, const double *e1) {
        const double dx2 = dX * dX, dy2 = dY * dY, dxy = dX * dY;
#pragma omp parallel for
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                int m = ix + Nx * iy;
                auto einv = [&](int jx, int jy) {
                    int q = jx + Nx * jy;
                    return (1. / e0[12 * (size_t)q + 2] + 1. / e1[12 * (size_t)q + 2]) / 2;
                };
                int pxx = (ix == Nx - 1) ? 1 : ix + 1, mxx = (ix == 0) ? Nx - 2 : ix - 1;
                int pyy = (iy == Ny - 1) ? 1 : iy + 1, myy = (iy == 0) ? Ny - 2 : iy - 1;
                double e = einv(ix, iy), epx = 0, emx = 0, epy = 0, emy = 0;
                if (Nx > 3) { epx = einv(pxx, iy); emx = einv(mxx, iy); }
                if (Ny > 3) { epy = einv(ix, pyy); emy = einv(ix, myy); }
                if (ix < Nx - 1 && Nx > 3) {
                    double cm = 0.5 * (emx + e) / dx2, cp = 0.5 * (epx + e) / dx2;
                    D1[TX][ci(TX, 0, 0, m)] = cm;
                    D1[TX][ci(TX, 0, 2, m)] = cp;
                    D1[TX][ci(TX, 0, 1, m)] = -cm - cp;
                }
                if (ix < Nx - 1 && iy < Ny - 1 && Nx > 3 && Ny > 3) {
                    double pp = 0.25 * epx / dxy, mm = 0.25 * emx / dxy;
                    D1[TXY][ci(TXY, 1, 3, m)] = pp;
                    D1[TXY][ci(TXY, 1, 0, m)] = mm;
                    D1[TXY][ci(TXY, 1, 2, m)] = -mm;
                    D1[TXY][ci(TXY, 1, 1, m)] = -pp;
                    double qp = 0.25 * epy / dxy, qm = 0.25 * emy / dxy;
                    D1[TXY][ci(TXY, 2, 3, m)] = qp;
                    D1[TXY][ci(TXY, 2, 0, m)] = qm;
                    D1[TXY][ci(TXY, 2, 2, m)] = -qp;
                    D1[TXY][ci(TXY, 2, 1, m)] = -qm;
                }
                if (iy < Ny - 1 && Ny > 3) {
                    double cp = 0.5 * (epy + e) / dy2, cm = 0.5 * (emy + e) / dy2;
                    D1[TY][ci(TY, 3, 2, m)] = cp;
                    D1[TY][ci(TY, 3, 0, m)] = cm;
                    D1[TY][ci(TY, 3, 1, m)] = -cm - cp;
                }
            }
    }

    // ParaxialPrimaryEvolutionOperator::update_D (:359-419): D = D0 + D1*K
    void update_D(const std::vector<Slab> &sl) {
        for (int t = 0; t < 3; t++) {
            const int nst = ns(t);
#pragma omp parallel for
            for (int iy = 0; iy < Ny; iy++)
                for (int ix = 0; ix < Nx; ix++) {
                    int m = ix + Nx * iy;
                    for (int b = 0; b < 4; b++)
                        for (int s = 0; s < nst; s++) D[t][ci(t, b, s, m)] = D0[t][ci(t, b, s, m)];
                    if (!inrow(t, ix, iy)) continue;
                    for (int s = 0; s < nst; s++) {
                        const Slab &k = sl[col(t, s, ix, iy)];
                        for (int rb = 0; rb < 2; rb++) {
                            double d10 = D1[t][ci(t, rb * 2 + 0, s, m)];
                            double d11 = D1[t][ci(t, rb * 2 + 1, s, m)];
                            // left block column: D1(.,c=0)*Kxx then D1(.,c=1)*Kxy
                            double &l = D[t][ci(t, rb * 2 + 0, s, m)];
                            l += d10 * k.kxx;
                            l += d11 * k.kxy;
                            double &r = D[t][ci(t, rb * 2 + 1, s, m)];
                            r += d10 * k.kxy;
                            r += d11 * k.kyy;
                        }
                    }
                }
        }
    }

    // ParaxialPrimaryEvolutionOperator::update (:212-357):
    // R = i W + (D Sinv + Sinv D)/4, W = (v_a d_b + d_a v_b)/2
    void update_R(const std::vector<Slab> &sl) {
        const double ddx = 0.5 / dX, ddy = 0.5 / dY;
        auto sinv = [](const Slab &s, double &ixx, double &iyy, double &ixy) {
            double det = s.sxx * s.syy - s.sxy * s.sxy;
            ixx = s.syy / det; iyy = s.sxx / det; ixy = -s.sxy / det;
        };
        for (int t = 0; t < 3; t++) {
            const int nst = ns(t);
#pragma omp parallel for
            for (int iy = 0; iy < Ny; iy++)
                for (int ix = 0; ix < Nx; ix++) {
                    int m = ix + Nx * iy;
                    for (int b = 0; b < 4; b++)
                        for (int s = 0; s < nst; s++) R[t][ci(t, b, s, m)] = 0;
                    if (!inrow(t, ix, iy)) continue;
                    const Slab &r = sl[m];
                    cd vr[2] = {r.vx * cd(0, 0.25), r.vy * cd(0, 0.25)};
                    // walk-off part: row-wise v_a(row) d_b and column-wise d_a v_b(col)
                    if (t != TXY) {
                        for (int s = 0; s < 3; s += 2) {
                            double d = (t == TX ? ddx : ddy) * (s == 0 ? -1 : 1);
                            const Slab &c = sl[col(t, s, ix, iy)];
                            cd vc[2] = {c.vx * cd(0, 0.25), c.vy * cd(0, 0.25)};
                            int bcol = (t == TX) ? 0 : 1;  // derivative direction
                            for (int a = 0; a < 2; a++) R[t][ci(t, a * 2 + bcol, s, m)] += d * vr[a];
                            for (int bb = 0; bb < 2; bb++) R[t][ci(t, bcol * 2 + bb, s, m)] += d * vc[bb];
                        }
                    }
                    // 0.25 * D * Sinv, Sinv taken at the column point
                    for (int s = 0; s < nst; s++) {
                        double ixx, iyy, ixy;
                        sinv(sl[col(t, s, ix, iy)], ixx, iyy, ixy);
                        for (int rb = 0; rb < 2; rb++) {
                            double d0 = D[t][ci(t, rb * 2 + 0, s, m)], d1 = D[t][ci(t, rb * 2 + 1, s, m)];
                            cd &l = R[t][ci(t, rb * 2 + 0, s, m)];
                            l += 0.25 * d0 * ixx;
                            l += 0.25 * d1 * ixy;
                            cd &rr = R[t][ci(t, rb * 2 + 1, s, m)];
                            rr += 0.25 * d0 * ixy;
                            rr += 0.25 * d1 * iyy;
                        }
                    }
                    // 0.25 * Sinv * D, Sinv taken at the row point
                    double ixx, iyy, ixy;
                    sinv(r, ixx, iyy, ixy);
                    for (int s = 0; s < nst; s++)
                        for (int cb = 0; cb < 2; cb++) {
                            double d0 = D[t][ci(t, 0 * 2 + cb, s, m)], d1 = D[t][ci(t, 1 * 2 + cb, s, m)];
                            cd &u = R[t][ci(t, 0 * 2 + cb, s, m)];
                            u += 0.25 * ixx * d0;
                            u += 0.25 * ixy * d1;
                            cd &w = R[t][ci(t, 1 * 2 + cb, s, m)];
                            w += 0.25 * ixy * d0;
                            w += 0.25 * iyy * d1;
                        }
                }
        }
    }

    // PhaseEvolutionOperator::update (:34-63)
    void update_phase(const double *e0, const double *e1) {
        const cd I(0, 1);
#pragma omp parallel for
        for (int m = 0; m < N; m++) {
            double w = 0.5, wpz = 0.5;
            if (p.weights) {
                size_t v = (size_t)m + (size_t)N * iz;
                w = p.weights[2 * v];
                wpz = p.weights[2 * v + 1];
            }
            const double *a = e0 + 12 * (size_t)m, *b = e1 + 12 * (size_t)m;
            double dxx = 0.5 * dZ * (w * a[9] + wpz * b[9] - no);
            double dyy = 0.5 * dZ * (w * a[10] + wpz * b[10] - no);
            double dxy = 0.5 * dZ * (w * a[11] + wpz * b[11]);
            double diff = (dxx - dyy) / 2., tr = (dxx + dyy) / 2.;
            double det = std::sqrt(dxy * dxy + diff * diff);
            cd sinc = (std::abs(det) < 1e-8) ? cd(1, 0) : std::sin(absorp * det) / det;
            cd ph = std::exp(I * absorp * tr), cs = std::cos(absorp * det);
            opxx[m] = ph * (cs + I * diff * sinc);
            opyy[m] = ph * (cs - I * diff * sinc);
            opxy[m] = ph * I * dxy * sinc;
        }
    }

    void update(int z) {
        iz = z;
        const double *e0 = eps_plane(z);
        const double *e1 = eps_plane(z + 1);
        std::vector<Slab> sl(N);
#pragma omp parallel for
        for (int m = 0; m < N; m++) sl[m] = slab_at(e0, e1, m);
        update_D1(e0, e1);
        update_D(sl);
        update_R(sl);
        update_phase(e0, e1);
    }

    // --- field operations on 2*N complex vectors -------------------------------
    // BlockSparseMatrix{DX,DY,DXDY}::block_vmult (BPL/BlockSparseMatrix.cpp:120-153,
    // 311-344, 509-550), including the duplicate-point copies
    void block_vmult(int t, int b, cd *dst, int dc, const cd *src, int sc, bool add) const {
        cd *d = dst + (size_t)dc * N;
        const cd *s = src + (size_t)sc * N;
        bool active = (t == TX) ? xper() : (t == TY) ? yper() : (xper() && yper());
        if (!active) {
            if (!add) std::fill(d, d + N, cd(0, 0));
            return;
        }
        const int nst = ns(t);
#pragma omp parallel for
        for (int iy = 0; iy < Ny; iy++)
            for (int ix = 0; ix < Nx; ix++) {
                int m = ix + Nx * iy;
                cd acc(0, 0);
                if (inrow(t, ix, iy))
                    for (int k = 0; k < nst; k++) acc += R[t][ci(t, b, k, m)] * s[col(t, k, ix, iy)];
                d[m] = add ? d[m] + acc : acc;
            }
        if (t == TX) {
            for (int iy = 0; iy < Ny; iy++) d[Nx - 1 + Nx * iy] = d[Nx * iy];
        } else if (t == TY) {
            for (int ix = 0; ix < Nx; ix++) d[ix + Nx * (Ny - 1)] = d[ix];
        } else {
            for (int iy = 1; iy < Ny - 1; iy++) d[Nx - 1 + Nx * iy] = d[Nx * iy];
            d[Nx - 1] = d[0];
            for (int ix = 0; ix < Nx; ix++) d[ix + Nx * (Ny - 1)] = d[ix];
        }
    }

    // BlockSparseMatrixDX/DY::shifted_block_vmult_inv (BPL/BlockSparseMatrix.cpp:194-258,
    // 385-450): dst = (I + s*A_block)^-1 src with s = -mu, per line, then duplicate <- first
    void block_solve(int dir, int b, cd *dst, int dc, const cd *src, int sc) const {
        cd *d = dst + (size_t)dc * N;
        const cd *s = src + (size_t)sc * N;
        const int L = dir == 0 ? Nx : Ny;
        if (L <= 3) {
            std::copy(s, s + N, d);
            return;
        }
        const int lines = dir == 0 ? Ny : Nx;
        const int t = dir == 0 ? TX : TY;
        const cd sf = -mu;
#pragma omp parallel for
        for (int l = 0; l < lines; l++) {
            auto at = [&](int j) { return dir == 0 ? (j + Nx * l) : (l + Nx * j); };
            auto sup = [&](int j) { return sf * R[t][ci(t, b, 2, at(j))]; };
            auto mid = [&](int j) { return 1. + sf * R[t][ci(t, b, 1, at(j))]; };
            auto inf = [&](int j) { return sf * R[t][ci(t, b, 0, at(j))]; };
            const int M = L - 1;
            std::vector<cd> gam(M - 1), corr(M);
            cd beta = 2. * mid(0);
            d[at(0)] = s[at(0)] / beta;
            corr[0] = -0.5;
            for (int j = 1; j < M - 1; j++) {
                gam[j - 1] = sup(j - 1) / beta;
                beta = mid(j) - inf(j) * gam[j - 1];
                d[at(j)] = (s[at(j)] - inf(j) * d[at(j - 1)]) / beta;
                corr[j] = -inf(j) * corr[j - 1] / beta;
            }
            gam[M - 2] = sup(M - 2) / beta;
            beta = mid(M - 1) + sup(M - 1) * inf(0) / mid(0) - inf(M - 1) * gam[M - 2];
            d[at(M - 1)] = (s[at(M - 1)] - inf(M - 1) * d[at(M - 2)]) / beta;
            corr[M - 1] = (sup(M - 1) - inf(M - 1) * corr[M - 2]) / beta;
            for (int j = M - 1; j > 0; j--) {
                d[at(j - 1)] -= gam[j - 1] * d[at(j)];
                corr[j - 1] -= gam[j - 1] * corr[j];
            }
            cd frac = (d[at(0)] - inf(0) / mid(0) * d[at(M - 1)]) /
                      (1. + corr[0] - inf(0) / mid(0) * corr[M - 1]);
            for (int j = 0; j < M; j++) d[at(j)] -= frac * corr[j];
            d[at(M)] = d[at(0)];
        }
    }

    void add_scaled(cd *dst, cd sfac, const cd *src) const {
#pragma omp parallel for
        for (int i = 0; i < 2 * N; i++) dst[i] += sfac * src[i];
    }
    void copy_comp(cd *dst, const cd *src, int c) const {
        std::copy(src + (size_t)c * N, src + (size_t)(c + 1) * N, dst + (size_t)c * N);
    }
    void add_comp(cd *dst, const cd *src, int c) const {
        for (int i = 0; i < N; i++) dst[(size_t)c * N + i] += src[(size_t)c * N + i];
    }
    void scale_comp(cd *dst, cd f, int c) const {
        for (int i = 0; i < N; i++) dst[(size_t)c * N + i] *= f;
    }

    // ParaxialPrimaryEvolutionOperator::apply (:153-209)
    int apply_primary(cd *f) {
        cd *t1 = tmp.data(), *t2 = tmp2.data();
        // forward y half-step
        block_vmult(TY, 0, t1, 0, f, 0, false);
        block_vmult(TY, 1, t1, 0, f, 1, true);
        block_vmult(TY, 3, t1, 1, f, 1, false);
        block_vmult(TXY, 3, t1, 1, f, 1, true);
        block_vmult(TY, 2, t1, 1, f, 0, true);
        block_vmult(TXY, 2, t1, 1, f, 0, true);
        add_scaled(f, mu, t1);
        // backward x half-step with Neumann iterations
        block_solve(0, 0, t1, 0, f, 0);
        block_solve(0, 3, t1, 1, f, 1);
        copy_comp(f, t1, 0);
        copy_comp(f, t1, 1);
        for (int it = 0; it < p.n_woodbury; it++) {
            block_vmult(TX, 1, t2, 0, t1, 1, false);
            block_vmult(TXY, 1, t2, 0, t1, 1, true);
            block_vmult(TXY, 0, t2, 0, t1, 0, true);
            block_vmult(TX, 2, t2, 1, t1, 0, false);
            scale_comp(t2, mu, 0);
            scale_comp(t2, mu, 1);
            block_solve(0, 0, t1, 0, t2, 0);
            block_solve(0, 3, t1, 1, t2, 1);
            add_comp(f, t1, 0);
            add_comp(f, t1, 1);
        }
        // forward x half-step
        block_vmult(TX, 0, t1, 0, f, 0, false);
        block_vmult(TXY, 0, t1, 0, f, 0, true);
        block_vmult(TX, 1, t1, 0, f, 1, true);
        block_vmult(TXY, 1, t1, 0, f, 1, true);
        block_vmult(TX, 3, t1, 1, f, 1, false);
        block_vmult(TX, 2, t1, 1, f, 0, true);
        add_scaled(f, mu, t1);
        // backward y half-step with Neumann iterations
        block_solve(1, 0, t1, 0, f, 0);
        block_solve(1, 3, t1, 1, f, 1);
        copy_comp(f, t1, 0);
        copy_comp(f, t1, 1);
        for (int it = 0; it < p.n_woodbury; it++) {
            block_vmult(TY, 1, t2, 0, t1, 1, false);
            block_vmult(TY, 2, t2, 1, t1, 0, false);
            block_vmult(TXY, 2, t2, 1, t1, 0, true);
            block_vmult(TXY, 3, t2, 1, t1, 1, true);
            scale_comp(t2, mu, 0);
            scale_comp(t2, mu, 1);
            block_solve(1, 0, t1, 0, t2, 0);
            block_solve(1, 3, t1, 1, t2, 1);
            add_comp(f, t1, 0);
            add_comp(f, t1, 1);
        }
        return 2 * p.n_woodbury;
    }

    // PhaseEvolutionOperator::apply (:14-32)
    void apply_phase(cd *f) const {
#pragma omp parallel for
        for (int m = 0; m < N; m++) {
            cd x = opxx[m] * f[m] + opxy[m] * f[m + N];
            cd y = opxy[m] * f[m] + opyy[m] * f[m + N];
            f[m] = x;
            f[m + N] = y;
        }
    }

    // SimpleShiftOperator (BPL/SimpleShiftOperator.cpp:3-93, wraps .h:24-39)
    struct Shift {
        std::vector<int> sx[4], sy[4];
        double wx[4] = {0, 0, 0, 0}, wy[4] = {0, 0, 0, 0};
        cd phase;
    };
    static void cr_weights(double nu, double *w) {
        w[0] = -nu * ((1 - nu) * (1 - nu)) / 2;
        w[1] = (1 - nu) * (1 + nu - 3 * nu * nu / 2);
        w[2] = nu * (1 + 4 * nu - 3 * nu * nu) / 2;
        w[3] = -nu * nu * (1 - nu) / 2;
    }
    static int wrap(int i, int n) {
        if (i > n - 2) return i % (n - 1);
        if (i < 0) return i + (1 - i / (n - 1)) * (n - 1);
        return i;
    }
    Shift make_shift(int iq) const {
        Shift sh;
        double qx = p.q[2 * iq], qy = p.q[2 * iq + 1];
        sh.phase = std::exp(-cd(0, dZ * (qx * qx + qy * qy) / (4 * nref)));
        if (Nx > 3) {
            double xs = -dZ * qx / (2 * nref * dX);
            double ms = std::floor(xs), nu = xs - ms;
            for (int k = 0; k < 4; k++)
                for (int ix = 0; ix < Nx; ix++) sh.sx[k].push_back(wrap((int)(ix + ms - 1 + k), Nx));
            cr_weights(nu, sh.wx);
        }
        if (Ny > 3) {
            double ys = -dZ * qy / (2 * nref * dX);  // deltaX, as the reference (:34)
            double ms = std::floor(ys), nu = ys - ms;
            for (int k = 0; k < 4; k++)
                for (int iy = 0; iy < Ny; iy++) sh.sy[k].push_back(wrap((int)(iy + ms - 1 + k), Ny));
            cr_weights(nu, sh.wy);
        }
        return sh;
    }
    void apply_shift(const Shift &sh, cd *f) {
        cd *t = tmp.data();
        if (Nx > 3) {
#pragma omp parallel for
            for (int iy = 0; iy < Ny; iy++)
                for (int c = 0; c < 2; c++)
                    for (int ix = 0; ix < Nx; ix++) {
                        cd acc(0, 0);
                        for (int k = 0; k < 4; k++) acc += sh.wx[k] * f[sh.sx[k][ix] + Nx * iy + (size_t)c * N];
                        t[ix + Nx * iy + (size_t)c * N] = acc;
                    }
        } else {
            std::copy(f, f + 2 * (size_t)N, t);
        }
        if (Ny > 3) {
#pragma omp parallel for
            for (int ix = 0; ix < Nx; ix++)
                for (int c = 0; c < 2; c++)
                    for (int iy = 0; iy < Ny; iy++) {
                        cd acc(0, 0);
                        for (int k = 0; k < 4; k++) acc += sh.wy[k] * t[ix + Nx * sh.sy[k][iy] + (size_t)c * N];
                        f[ix + Nx * iy + (size_t)c * N] = acc;
                    }
        } else {
            std::copy(t, t + 2 * (size_t)N, f);
        }
#pragma omp parallel for
        for (int i = 0; i < 2 * N; i++) f[i] *= sh.phase;
    }

    // Input/OutputFresnelOperator::apply (BPL/FresnelOperator.cpp:16-38, 45-69)
    void fresnel_in(cd *f, double niso) {
        const double *e = eps_plane(0);
#pragma omp parallel for
        for (int m = 0; m < N; m++) {
            double nxx = e[12 * (size_t)m + 9], nyy = e[12 * (size_t)m + 10], nxy = e[12 * (size_t)m + 11];
            double det = (niso + nxx) * (niso + nyy) - nxy * nxy;
            cd a = f[m], b = f[m + N];
            f[m] = 2 * niso * (niso + nyy) / det * a - 2 * niso * nxy / det * b;
            f[m + N] = -2 * niso * nxy / det * a + 2 * niso * (niso + nxx) / det * b;
        }
    }
    void fresnel_out(cd *f, double niso) {
        const double *e = eps_plane(Nz - 1);
#pragma omp parallel for
        for (int m = 0; m < N; m++) {
            double nxx = e[12 * (size_t)m + 9], nyy = e[12 * (size_t)m + 10], nxy = e[12 * (size_t)m + 11];
            double det = (niso + nxx) * (niso + nyy) - nxy * nxy;
            cd a = f[m], b = f[m + N];
            f[m] = 2. * ((1. - niso * (niso + nyy) / det) * a + niso * nxy / det * b);
            f[m + N] = 2. * ((1. - niso * (niso + nxx) / det) * b + niso * nxy / det * a);
        }
    }

    // ScreenOutput::assemble_fourier_filter (PP/ScreenOutput.cpp:226-302) and the
    // filter loop of apply_no_export (:197-223), with nout(lambda) per wavelength
    void filter_wavelength(int w, cd *Ewl) {
        const double wl = p.wl[w];
        const double kk0 = 2 * kPi / wl;
        std::vector<double> idx(p.iso_index, p.iso_index + p.n_iso);
        idx.push_back(p.nout[w]);
        double zf = p.z_foc;
        for (int i = 0; i < p.n_iso; i++) zf -= p.iso_thick[i] * p.nout_ref / p.iso_index[i];
        const double n_out = idx.back();
        fft2_planes(Nx, Ny, 4 * p.n_q, -1, Ewl);
        std::vector<cd> F(4 * (size_t)N);
        for (int iq = 0; iq < p.n_q; iq++) {
            const double qx0 = p.q[2 * iq], qy0 = p.q[2 * iq + 1];
#pragma omp parallel for
            for (int iy = 0; iy < Ny; iy++)
                for (int ix = 0; ix < Nx; ix++) {
                    double kx = (ix < double(Nx / 2.)) ? 2 * kPi * ix / (p.dx * Nx) : -2 * kPi * (Nx - ix) / (p.dx * Nx);
                    double ky = (iy < double(Ny / 2.)) ? 2 * kPi * iy / (p.dy * Ny) : -2 * kPi * (Ny - iy) / (p.dy * Ny);
                    double q0 = kx / kk0 + qx0, q1 = ky / kk0 + qy0;
                    double qn = std::sqrt(q0 * q0 + q1 * q1);
                    double qns = qn * qn;
                    double q0n = std::sqrt(qx0 * qx0 + qy0 * qy0);  // |q0|, as the reference (:252)
                    cd *Fm = &F[4 * ((size_t)ix + (size_t)Nx * iy)];
                    if (qn < p.NA) {
                        cd T[4][4];
                        for (int i = 0; i < 4; i++)
                            for (int j = 0; j < 4; j++) T[i][j] = (i == j) ? 1. : 0.;
                        for (int l = 0; l < p.n_iso; l++) {
                            double np = idx[l], nn = idx[l + 1];
                            double neff = std::sqrt(np * np - qns), nneff = std::sqrt(nn * nn - qns);
                            double im[2][2] = {{neff + q0 * q0 / neff, q0 * q1 / neff},
                                               {q0 * q1 / neff, neff + q1 * q1 / neff}};
                            double ni[2][2] = {{(1 - (q0 / nn) * (q0 / nn)) / nneff, -q0 * q1 / (nn * nn * nneff)},
                                               {-q0 * q1 / (nn * nn * nneff), (1 - (q1 / nn) * (q1 / nn)) / nneff}};
                            double tp[2][2], tm[2][2];
                            for (int i = 0; i < 2; i++)
                                for (int j = 0; j < 2; j++) {
                                    double prod = ni[i][0] * im[0][j] + ni[i][1] * im[1][j];
                                    tp[i][j] = 0.5 * ((i == j ? 1. : 0.) + prod);
                                    tm[i][j] = (i == j ? 1. : 0.) - tp[i][j];
                                }
                            cd g = std::exp(cd(0, (neff - std::sqrt(np * np - q0n)) * kk0 * p.iso_thick[l]));
                            cd Lm[4][4];
                            for (int i = 0; i < 2; i++)
                                for (int j = 0; j < 2; j++) {
                                    Lm[i][j] = tp[i][j] * g;
                                    Lm[i][j + 2] = tm[i][j] * std::conj(g);
                                    Lm[i + 2][j] = tm[i][j] * g;
                                    Lm[i + 2][j + 2] = tp[i][j] * std::conj(g);
                                }
                            cd Tn[4][4];
                            for (int i = 0; i < 4; i++)
                                for (int j = 0; j < 4; j++) {
                                    cd acc(0, 0);
                                    for (int k = 0; k < 4; k++) acc += Lm[i][k] * T[k][j];
                                    Tn[i][j] = acc;
                                }
                            for (int i = 0; i < 4; i++)
                                for (int j = 0; j < 4; j++) T[i][j] = Tn[i][j];
                        }
                        cd det = T[2][2] * T[3][3] - T[2][3] * T[3][2];
                        cd inv[2][2] = {{T[3][3] / det, -T[2][3] / det}, {-T[3][2] / det, T[2][2] / det}};
                        cd ph = std::exp(cd(0, (std::sqrt(n_out * n_out - qns) - std::sqrt(n_out * n_out - q0n)) * kk0 * zf));
                        for (int i = 0; i < 2; i++)
                            for (int j = 0; j < 2; j++) {
                                cd acc = T[i][j];
                                for (int k = 0; k < 2; k++)
                                    for (int l2 = 0; l2 < 2; l2++) acc -= T[i][k + 2] * inv[k][l2] * T[l2 + 2][j];
                                Fm[2 * i + j] = acc * ph;
                            }
                    } else {
                        for (int i = 0; i < 4; i++) Fm[i] = 0;
                    }
                }
            for (int pol = 0; pol < 2; pol++) {
                cd *f = Ewl + (size_t)(2 * iq + pol) * 2 * N;
#pragma omp parallel for
                for (int m = 0; m < N; m++) {
                    const cd *Fm = &F[4 * (size_t)m];
                    cd a = Fm[0] * f[m] + Fm[1] * f[m + N];
                    cd b = Fm[2] * f[m] + Fm[3] * f[m + N];
                    f[m] = a / double(Nx * Ny);
                    f[m + N] = b / double(Nx * Ny);
                }
            }
        }
        fft2_planes(Nx, Ny, 4 * p.n_q, +1, Ewl);
    }

    // BPMIteration::propagate_fields (BPL/BPMIteration.cpp:53-177) + screen filter
    // bulk (optional): the field of every transverse plane, saved after the
    // input interface (plane 0) and after each step (plane z+1), before the
    // output interface (BPMIteration.cpp:111-120, 153-168); layout
    // [w][q][pol][plane][comp][m]
    int propagate(cd *E, cd *bulk = nullptr) {
        const size_t stride = 2 * (size_t)N;
        auto save = [&](int w, int iq, int pol, int plane, const cd *f) {
            if (!bulk) return;
            cd *b = bulk + ((((size_t)w * p.n_q + iq) * 2 + pol) * Nz + plane) * stride;
            std::memcpy(b, f, sizeof(cd) * stride);
        };
        for (int w = 0; w < p.n_wl; w++) {
            set_wavelength(w);
            std::vector<Shift> shifts;
            for (int iq = 0; iq < p.n_q; iq++) shifts.push_back(make_shift(iq));
            double nout_f = p.n_iso > 0 ? p.iso_index[0] : p.nout[w];
            for (int pol = 0; pol < 2; pol++)
                for (int iq = 0; iq < p.n_q; iq++) {
                    cd *f = E + (size_t)(pol + 2 * (iq + p.n_q * w)) * stride;
                    if (p.beam == 2) continue;
                    double th = 0.5 * kPi * pol;
                    for (int m = 0; m < N; m++) {
                        double x = ((m % Nx) - (Nx - 1) / 2.) * p.dx;
                        double y = ((m / Nx) - (Ny - 1) / 2.) * p.dy;
                        double g = (p.beam == 1) ? std::exp(-(x * x + y * y) / (2 * p.waist * p.waist)) : 1.0;
                        if (p.beam == 1) {
                            f[m] = std::cos(th) * g;
                            f[m + N] = std::sin(th) * g;
                        } else {
                            f[m] = std::cos(th);
                            f[m + N] = std::sin(th);
                        }
                    }
                    fresnel_in(f, p.nin[w]);
                }
            for (int pol = 0; pol < 2; pol++)
                for (int iq = 0; iq < p.n_q; iq++)
                    save(w, iq, pol, 0, E + (size_t)(pol + 2 * (iq + p.n_q * w)) * stride);
            for (int z = 0; z < Nz - 1; z++) {
                update(z);
                for (int pol = 0; pol < 2; pol++)
                    for (int iq = 0; iq < p.n_q; iq++) {
                        cd *f = E + (size_t)(pol + 2 * (iq + p.n_q * w)) * stride;
                        apply_shift(shifts[iq], f);
                        apply_phase(f);
                        apply_primary(f);
                        apply_phase(f);
                        apply_shift(shifts[iq], f);
                        save(w, iq, pol, z + 1, f);
                    }
            }
            for (int pol = 0; pol < 2; pol++)
                for (int iq = 0; iq < p.n_q; iq++)
                    fresnel_out(E + (size_t)(pol + 2 * (iq + p.n_q * w)) * stride, nout_f);
        }
        if (p.apply_filter)
            for (int w = 0; w < p.n_wl; w++) filter_wavelength(w, E + (size_t)w * p.n_q * 2 * stride);
        return 0;
    }
};

int validate(const OracleParams *p) {
    if (!p || p->Nx < 1 || p->Ny < 1 || p->Nz < 1) return -1;
    if (p->field_dim != 3 && p->field_dim != 6) return -2;
    if (p->n_wl < 1 || p->n_q < 1) return -3;
    return 0;
}

struct ThreadScope {
    int saved;
    explicit ThreadScope(int n) : saved(omp_get_max_threads()) {
        omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
    }
    ~ThreadScope() { omp_set_num_threads(saved); }
};

}  // namespace

extern "C" {

int oracle_propagate(const OracleParams *p, const double *lc, double *E) {
    int rc = validate(p);
    if (rc) return rc;
    ThreadScope ts(p->nthreads);
    Oracle o(p, lc);
    return o.propagate(reinterpret_cast<cd *>(E));
}

int oracle_propagate_bulk(const OracleParams *p, const double *lc, double *E, double *bulk) {
    int rc = validate(p);
    if (rc) return rc;
    ThreadScope ts(p->nthreads);
    Oracle o(p, lc);
    return o.propagate(reinterpret_cast<cd *>(E), reinterpret_cast<cd *>(bulk));
}

void *oracle_create(const OracleParams *p, const double *lc) {
    if (validate(p)) return nullptr;
    return new Oracle(p, lc);
}
void oracle_destroy(void *h) { delete static_cast<Oracle *>(h); }
int oracle_set_wavelength(void *h, int iwl) {
    Oracle *o = static_cast<Oracle *>(h);
    ThreadScope ts(o->p.nthreads);
    o->set_wavelength(iwl);
    return 0;
}
double oracle_nref(void *h) { return static_cast<Oracle *>(h)->nref; }
void oracle_eps_plane(void *h, int iz, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    const double *e = o->eps_plane(iz);
    std::memcpy(out, e, sizeof(double) * 12 * (size_t)o->N);
}
void oracle_update(void *h, int iz) {
    Oracle *o = static_cast<Oracle *>(h);
    ThreadScope ts(o->p.nthreads);
    o->update(iz);
}
void oracle_get_R(void *h, int type, int block, int shift, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    cd *d = reinterpret_cast<cd *>(out);
    for (int m = 0; m < o->N; m++) d[m] = o->R[type][o->ci(type, block, shift, m)];
}
void oracle_block_vmult(void *h, int type, int block, double *dst, int dcomp, const double *src, int scomp, int add) {
    Oracle *o = static_cast<Oracle *>(h);
    o->block_vmult(type, block, reinterpret_cast<cd *>(dst), dcomp, reinterpret_cast<const cd *>(src), scomp, add != 0);
}
void oracle_block_solve(void *h, int dir, int block, double *dst, int dcomp, const double *src, int scomp) {
    Oracle *o = static_cast<Oracle *>(h);
    o->block_solve(dir, block, reinterpret_cast<cd *>(dst), dcomp, reinterpret_cast<const cd *>(src), scomp);
}
void oracle_cyclic_line_solve(int M, const double *a, const double *b, const double *c, const double *d, double *x) {
    // same recurrence as the block solver, on explicit coefficient arrays
    const cd *A = reinterpret_cast<const cd *>(a), *B = reinterpret_cast<const cd *>(b);
    const cd *C = reinterpret_cast<const cd *>(c), *Dv = reinterpret_cast<const cd *>(d);
    cd *X = reinterpret_cast<cd *>(x);
    std::vector<cd> gam(M > 1 ? M - 1 : 1), corr(M);
    cd beta = 2. * B[0];
    X[0] = Dv[0] / beta;
    corr[0] = -0.5;
    for (int j = 1; j < M - 1; j++) {
        gam[j - 1] = C[j - 1] / beta;
        beta = B[j] - A[j] * gam[j - 1];
        X[j] = (Dv[j] - A[j] * X[j - 1]) / beta;
        corr[j] = -A[j] * corr[j - 1] / beta;
    }
    gam[M - 2] = C[M - 2] / beta;
    beta = B[M - 1] + C[M - 1] * A[0] / B[0] - A[M - 1] * gam[M - 2];
    X[M - 1] = (Dv[M - 1] - A[M - 1] * X[M - 2]) / beta;
    corr[M - 1] = (C[M - 1] - A[M - 1] * corr[M - 2]) / beta;
    for (int j = M - 1; j > 0; j--) {
        X[j - 1] -= gam[j - 1] * X[j];
        corr[j - 1] -= gam[j - 1] * corr[j];
    }
    cd frac = (X[0] - A[0] / B[0] * X[M - 1]) / (1. + corr[0] - A[0] / B[0] * corr[M - 1]);
    for (int j = 0; j < M; j++) X[j] -= frac * corr[j];
}
int oracle_apply_primary(void *h, double *field) {
    Oracle *o = static_cast<Oracle *>(h);
    ThreadScope ts(o->p.nthreads);
    return o->apply_primary(reinterpret_cast<cd *>(field));
}
void oracle_apply_phase(void *h, double *field) {
    static_cast<Oracle *>(h)->apply_phase(reinterpret_cast<cd *>(field));
}
void oracle_apply_shift(void *h, int iq, double *field) {
    Oracle *o = static_cast<Oracle *>(h);
    o->apply_shift(o->make_shift(iq), reinterpret_cast<cd *>(field));
}
void oracle_phase_coefs(void *h, double *out) {
    Oracle *o = static_cast<Oracle *>(h);
    cd *d = reinterpret_cast<cd *>(out);
    for (int m = 0; m < o->N; m++) {
        d[m] = o->opxx[m];
        d[m + o->N] = o->opyy[m];
        d[m + 2 * (size_t)o->N] = o->opxy[m];
    }
}
void oracle_filter_wavelength(void *h, int iwl, double *E_wl) {
    Oracle *o = static_cast<Oracle *>(h);
    ThreadScope ts(o->p.nthreads);
    o->filter_wavelength(iwl, reinterpret_cast<cd *>(E_wl));
}
void oracle_fft2(int Nx, int Ny, int howmany, int sign, double *data) {
    fft2_planes(Nx, Ny, howmany, sign, reinterpret_cast<cd *>(data));
}

}  // extern "C"
