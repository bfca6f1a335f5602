// Host side of the microscope imaging step: geometry, q-integration weights,
// Jones matrices and the device buffers of one set of output fields.
#include "viewer.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "settings.h"

namespace nmk {

#define VIEWCHK(x)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

namespace {
constexpr double kPi = 3.141592653589793;

// number of wavevectors up to ring r (field_viewer.py:516-517)
int q_index(int r) { return 1 + 3 * r * (r - 1); }

// primitive used by the q-integration rule (field_viewer.py:514-515)
double q_int(double q, double qp) { return q - std::atanh(q) - qp * std::log(1 - q * q) / 2; }

template <class P>
void dev_alloc(P **p, size_t n) {
    VIEWCHK(hipMalloc((void **)p, std::max<size_t>(1, n) * sizeof(P)));
}
}  // namespace

void condenser_wavevectors(ViewGeometry &g) {
    const int Nr = g.Nr;
    g.qx.clear();
    g.qy.clear();
    auto beta = [&](int ir) { return ir * g.max_NA_condenser / (Nr - 1); };
    if ((g.Nx == 1 || g.Ny == 1) && g.koehler_1D) {
        // rays along the only transverse axis with more than one point
        const bool along_y = g.Nx == 1;
        g.qx.assign(2 * Nr - 1, 0.0);
        g.qy.assign(2 * Nr - 1, 0.0);
        std::vector<double> &a = along_y ? g.qy : g.qx;
        for (int ir = 1; ir < Nr; ir++) {
            a[Nr - 1 + ir] = beta(ir);
            a[Nr - 1 - ir] = -beta(ir);
        }
    } else {
        g.qx.assign(q_index(Nr), 0.0);
        g.qy.assign(q_index(Nr), 0.0);
        for (int ir = 1; ir < Nr; ir++)
            for (int iphi = 0; iphi < 6 * ir; iphi++) {
                const double phi = iphi * kPi / (3 * ir);
                g.qx[q_index(ir) + iphi] = beta(ir) * std::cos(phi);
                g.qy[q_index(ir) + iphi] = beta(ir) * std::sin(phi);
            }
    }
}

void validate_geometry(ViewGeometry &g) {
    if (g.Nx < 1 || g.Ny < 1) throw BpmError("mesh_dimensions should be positive");
    if (g.Nx > 4097 || g.Ny > 4097) throw BpmError("mesh dimensions above 4097 points per axis are not supported");
    if (g.wl.empty()) throw BpmError("at least one wavelength is required");
    for (double l : g.wl)
        if (!(l > 0)) throw BpmError("wavelengths must be positive");
    if (g.n_out.empty()) g.n_out.assign(g.wl.size(), 1.0);
    if (g.n_out.size() != g.wl.size()) throw BpmError("n_out needs one value per wavelength");
    if (g.Nr < 1) throw BpmError("N_radial_wavevectors must be >= 1");
    if (g.qx.empty()) condenser_wavevectors(g);
    if (g.qx.size() != g.qy.size() || g.qx.empty()) throw BpmError("wavevectors must be (qx, qy) pairs");
    const int Nq = (int)g.qx.size();
    const bool line = (g.Nx == 1 || g.Ny == 1) && g.koehler_1D;
    if (line ? Nq != 2 * g.Nr - 1 : Nq != q_index(g.Nr))
        throw BpmError("wrong number of wavevectors for N_radial_wavevectors = " + std::to_string(g.Nr));
}

QuadRule condenser_quadrature(const ViewGeometry &g, double NA) {
    QuadRule r;
    const int Nr = g.Nr;
    const int Nq = (int)g.qx.size();
    if (!(NA >= 0) || !std::isfinite(NA)) throw BpmError("NA_condenser must be a finite value >= 0");
    const double dq = (Nr > 1 && g.max_NA_condenser > 0) ? g.max_NA_condenser / (Nr - 1) : 1.0;
    const double rq = std::ceil(NA / dq);
    const int qr_max = rq >= Nr - 1 ? Nr - 1 : (int)rq;
    r.w.assign(Nq, 0.0);
    if (!g.koehler_1D) {
        r.q0 = 0;
        r.q1 = q_index(qr_max + 1);
    } else {
        r.q0 = Nr - 1 - qr_max;
        r.q1 = Nr + qr_max;
    }
    if (NA < 1e-8) {
        // only the normal-incidence wave (field_viewer.py:384-389)
        const int q = g.koehler_1D ? Nr - 1 : 0;
        r.w[q] = 1.0;
        r.q0 = q;
        r.q1 = q + 1;
        return r;
    }
    if (!g.koehler_1D) {
        // hat functions in q times g(q) = q/(1-q^2), integrated exactly
        // (field_viewer.py:401-428), times 2/NA^2 (:429-432)
        std::vector<double> iw(q_index(qr_max + 1), 0.0);
        if (qr_max == 1) {
            iw[0] = (q_int(NA, dq) - q_int(0, dq)) / dq;
            for (int q = 1; q < 7; q++) iw[q] = (q_int(0, 0) - q_int(NA, 0)) / (6 * dq);
        } else {
            iw[0] = (q_int(dq, dq) - q_int(0, dq)) / dq;
            for (int ir = 1; ir < qr_max - 1; ir++) {
                const double v = (q_int((ir + 1) * dq, (ir + 1) * dq) + q_int((ir - 1) * dq, (ir - 1) * dq) -
                                  2 * q_int(ir * dq, ir * dq)) /
                                 (6 * ir * dq);
                for (int q = q_index(ir); q < q_index(ir + 1); q++) iw[q] = v;
            }
            const int a = qr_max - 1;
            const double va = (q_int(NA, qr_max * dq) + q_int((qr_max - 2) * dq, (qr_max - 2) * dq) -
                               2 * q_int(a * dq, a * dq)) /
                              (6 * a * dq);
            for (int q = q_index(a); q < q_index(a + 1); q++) iw[q] = va;
            const double vb = (q_int(a * dq, a * dq) - q_int(NA, a * dq)) / (6 * qr_max * dq);
            for (int q = q_index(qr_max); q < q_index(qr_max + 1); q++) iw[q] = vb;
        }
        const double pre = 2 / (NA * NA);
        for (size_t q = 0; q < iw.size(); q++) r.w[q] = pre * iw[q];
    } else {
        // trapezoidal rule with partial edge cells (field_viewer.py:433-446)
        const double edge = NA - (qr_max - 1) * dq;
        std::vector<double> iw(2 * qr_max + 1, dq);
        const int last = 2 * qr_max;
        if (qr_max == 1) {
            iw[0] = iw[last] = edge * edge / (2 * dq);
            iw[1] = edge * (2 - edge / dq);
        } else {
            iw[0] = iw[last] = edge * edge / (2 * dq);
            iw[1] = iw[last - 1] = edge * (1 - edge / (2 * dq)) + dq / 2;
        }
        for (int k = 0; k <= last; k++) r.w[r.q0 + k] = iw[k] / (2 * NA);
    }
    return r;
}

namespace {
void identity(c64 *m) {
    m[0] = m[3] = mk(1.0, 0.0);
    m[1] = m[2] = mk(0.0, 0.0);
}
void mul(const c64 *a, const c64 *b, c64 *o) {
    c64 t[4] = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                a[2] * b[1] + a[3] * b[3]};
    for (int k = 0; k < 4; k++) o[k] = t[k];
}
// field_viewer.py:472-477
void polariser(double angle, c64 *m) {
    const double c = std::cos(angle * kPi / 180), s = std::sin(angle * kPi / 180);
    m[0] = mk(c * c, 0.0);
    m[1] = m[2] = mk(c * s, 0.0);
    m[3] = mk(s * s, 0.0);
}
// field_viewer.py:479-510
void waveplate(int kind, double angle, double wl, c64 *m) {
    const double c = std::cos(angle * kPi / 180), s = std::sin(angle * kPi / 180);
    if (kind == WP_QUARTER) {
        const c64 f = mk(std::cos(-kPi / 4), std::sin(-kPi / 4));
        m[0] = f * mk(c * c, s * s);
        m[1] = m[2] = f * mk(c * s, -c * s);
        m[3] = f * mk(s * s, c * c);
    } else if (kind == WP_HALF) {
        const c64 f = mk(std::cos(-kPi / 2), std::sin(-kPi / 2));
        m[0] = f * mk(c * c - s * s, 0.0);
        m[1] = m[2] = f * mk(2 * c * s, 0.0);
        m[3] = f * mk(s * s - c * c, 0.0);
    } else if (kind == WP_TINT) {
        // full wave at 540 nm for a wavelength-independent index
        const double ph = kPi * 0.540 / wl;
        const c64 g = mk(std::cos(ph), std::sin(ph)), gc = conj(g);
        m[0] = (c * c) * gc + (s * s) * g;
        m[1] = m[2] = (c * s) * (gc - g);
        m[3] = (s * s) * gc + (c * c) * g;
    } else {
        throw BpmError("unknown waveplate type " + std::to_string(kind));
    }
}
}  // namespace

std::vector<JonesSet> jones_matrices(const ViewGeometry &g, const ViewOptics &o) {
    std::vector<JonesSet> out(g.wl.size());
    for (size_t l = 0; l < g.wl.size(); l++) {
        JonesSet &J = out[l];
        identity(J.R);
        identity(J.Wu);
        identity(J.A);
        // element order of FieldViewer._update_specter (field_viewer.py:364-385)
        if (o.polariser) polariser(o.polariser_angle, J.R);
        if (o.lower_waveplate != WP_NONE) {
            c64 W[4];
            waveplate(o.lower_waveplate, o.lower_waveplate_angle, g.wl[l], W);
            if (o.polariser) mul(W, J.R, J.R);
            else for (int k = 0; k < 4; k++) J.R[k] = W[k];
        }
        if (o.upper_waveplate != WP_NONE) waveplate(o.upper_waveplate, o.upper_waveplate_angle, g.wl[l], J.Wu);
        if (o.analyser) polariser(o.analyser_angle, J.A);
    }
    return out;
}

Viewer::Viewer(const ViewGeometry &geo) : g_(geo) {
    validate_geometry(g_);
    if (g_.device >= 0) VIEWCHK(hipSetDevice(g_.device));
    Nl_ = (int)g_.wl.size();
    Nq_ = (int)g_.qx.size();
    N_ = (int64_t)g_.Nx * g_.Ny;
    dx_ = g_.Lx / std::max(1, g_.Nx - 1);
    dy_ = g_.Ly / std::max(1, g_.Ny - 1);
    NA_ = g_.max_NA_objective;
    px_ = make_fft_plan(g_.Nx);
    py_ = make_fft_plan(g_.Ny);
    try {
        VIEWCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        for (auto &e : ev_) VIEWCHK(hipEventCreate(&e));
        const size_t nv = (size_t)g_.n_values();
        dev_alloc(&vals_, nv);
        dev_alloc(&focused_buf_, nv);
        dev_alloc(&filt_, (size_t)Nl_ * Nq_ * N_);
        dev_alloc(&spec_, (size_t)Nl_ * N_);
        dev_alloc(&spec_out_, (size_t)Nl_ * N_);
        dev_alloc(&image_, (size_t)3 * N_);
        dev_alloc(&w_, (size_t)Nq_);
        dev_alloc(&cmf_, (size_t)3 * Nl_);
        dev_alloc(&jones_, (size_t)Nl_);
        dev_alloc(&planes_, (size_t)Nl_ * Nq_);
        dev_alloc(&twx_, (size_t)g_.Nx);
        dev_alloc(&twy_, (size_t)g_.Ny);
        VIEWCHK(hipMemsetAsync(vals_, 0, nv * sizeof(c64), st_));
        auto twiddles = [&](int n, c64 *d) {
            std::vector<c64> h(n);
            for (int k = 0; k < n; k++) {
                const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
                h[k] = mk((double)std::cos(a), (double)std::sin(a));
            }
            VIEWCHK(hipMemcpy(d, h.data(), n * sizeof(c64), hipMemcpyHostToDevice));
        };
        twiddles(g_.Nx, twx_);
        twiddles(g_.Ny, twy_);
        // k0 = 2 pi / lambda, kn = 2 pi n_out / lambda, p = k0 q
        // (light_propagator.py:608-611)
        std::vector<ViewPlane> pl((size_t)Nl_ * Nq_);
        for (int l = 0; l < Nl_; l++)
            for (int q = 0; q < Nq_; q++) {
                ViewPlane &p = pl[(size_t)l * Nq_ + q];
                p.k0 = 2 * kPi / g_.wl[l];
                p.kn = 2 * kPi * g_.n_out[l] / g_.wl[l];
                p.px = p.k0 * g_.qx[q];
                p.py = p.k0 * g_.qy[q];
            }
        VIEWCHK(hipMemcpy(planes_, pl.data(), pl.size() * sizeof(ViewPlane), hipMemcpyHostToDevice));
        VIEWCHK(hipStreamSynchronize(st_));
    } catch (...) {
        release();
        throw;
    }
}

Viewer::~Viewer() { release(); }

void Viewer::release() {
    if (st_) (void)hipStreamSynchronize(st_);
    for (void *p : {(void *)vals_, (void *)focused_buf_, (void *)filt_, (void *)spec_, (void *)spec_out_,
                    (void *)image_, (void *)w_, (void *)cmf_, (void *)jones_, (void *)planes_, (void *)twx_,
                    (void *)twy_})
        if (p) (void)hipFree(p);
    vals_ = focused_buf_ = filt_ = twx_ = twy_ = nullptr;
    spec_ = spec_out_ = image_ = w_ = cmf_ = nullptr;
    jones_ = nullptr;
    planes_ = nullptr;
    for (auto &e : ev_)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (st_) (void)hipStreamDestroy(st_);
    st_ = nullptr;
}

void Viewer::set_fields(const c64 *E) {
    VIEWCHK(hipMemcpyAsync(vals_, E, (size_t)g_.n_values() * sizeof(c64), hipMemcpyHostToDevice, st_));
    VIEWCHK(hipStreamSynchronize(st_));
    focused_ = has_spec_ = has_image_ = false;
}

void Viewer::set_fields_device(const c64 *E, hipStream_t producer) {
    if (producer) VIEWCHK(hipStreamSynchronize(producer));
    VIEWCHK(hipMemcpyAsync(vals_, E, (size_t)g_.n_values() * sizeof(c64), hipMemcpyDeviceToDevice, st_));
    VIEWCHK(hipStreamSynchronize(st_));
    focused_ = has_spec_ = has_image_ = false;
}

void Viewer::focus(double z, double NA) {
    // OpticalFields.update_NA_objective (light_propagator.py:691-693)
    if (!std::isnan(NA)) NA_ = std::max(0.0, std::min(g_.max_NA_objective, NA));
    if (!std::isnan(z)) z_ = z;
    // focus_fields (light_propagator.py:696-719)
    const double zf = z_ + NA_ * NA_ * g_.zfoc_NA_corr;
    const int nlq = Nl_ * Nq_;
    VIEWCHK(hipEventRecord(ev_[0], st_));
    launch_view_filter(filt_, g_.Nx, g_.Ny, dx_, dy_, planes_, nlq, NA_, zf, st_);
    VIEWCHK(hipEventRecord(ev_[1], st_));
    launch_fft_rows(vals_, focused_buf_, g_.Nx, g_.Ny, nlq * 4, -1, twx_, px_, st_);
    VIEWCHK(hipEventRecord(ev_[2], st_));
    launch_view_columns(focused_buf_, filt_, g_.Nx, g_.Ny, nlq, twy_, py_, st_);
    VIEWCHK(hipEventRecord(ev_[3], st_));
    launch_fft_rows(focused_buf_, focused_buf_, g_.Nx, g_.Ny, nlq * 4, +1, twx_, px_, st_);
    VIEWCHK(hipEventRecord(ev_[4], st_));
    VIEWCHK(hipEventSynchronize(ev_[4]));
    float t;
    for (int k = 0; k < 4; k++) {
        VIEWCHK(hipEventElapsedTime(&t, ev_[k], ev_[k + 1]));
        ms_[T_FILTER + k] = t;
    }
    focused_ = true;
    has_spec_ = has_image_ = false;
}

void Viewer::download_focused(c64 *E) {
    if (!focused_) throw BpmError("the fields have not been focused yet");
    VIEWCHK(hipMemcpyAsync(E, focused_buf_, (size_t)g_.n_values() * sizeof(c64), hipMemcpyDeviceToHost, st_));
    VIEWCHK(hipStreamSynchronize(st_));
}

void Viewer::render(const ViewOptics &o, const double *cmf) {
    if (!focused_) throw BpmError("the fields have not been focused yet");
    const QuadRule qr = condenser_quadrature(g_, o.NA_condenser);
    h_w_ = qr.w;
    h_jones_ = jones_matrices(g_, o);
    VIEWCHK(hipMemcpyAsync(w_, h_w_.data(), h_w_.size() * sizeof(double), hipMemcpyHostToDevice, st_));
    VIEWCHK(hipMemcpyAsync(jones_, h_jones_.data(), h_jones_.size() * sizeof(JonesSet), hipMemcpyHostToDevice, st_));
    if (cmf) {
        h_cmf_.assign(cmf, cmf + 3 * Nl_);
        VIEWCHK(hipMemcpyAsync(cmf_, h_cmf_.data(), h_cmf_.size() * sizeof(double), hipMemcpyHostToDevice, st_));
    }
    VIEWCHK(hipEventRecord(ev_[4], st_));
    launch_view_spectrum(focused_buf_, N_, Nl_, Nq_, qr.q0, qr.q1, w_, jones_, spec_, st_);
    VIEWCHK(hipEventRecord(ev_[5], st_));
    launch_view_colour(spec_, N_, Nl_, cmf_, o.intensity, o.grayscale ? 1 : 0, o.gamma ? 1 : 0, spec_out_,
                       cmf ? image_ : nullptr, st_);
    VIEWCHK(hipEventRecord(ev_[6], st_));
    VIEWCHK(hipEventSynchronize(ev_[6]));
    float t;
    VIEWCHK(hipEventElapsedTime(&t, ev_[4], ev_[5]));
    ms_[T_SPECTRUM] = t;
    VIEWCHK(hipEventElapsedTime(&t, ev_[5], ev_[6]));
    ms_[T_COLOUR] = t;
    has_spec_ = true;
    has_image_ = cmf != nullptr;
}

void Viewer::download_spectrum(double *out) {
    if (!has_spec_) throw BpmError("no spectrum has been computed yet");
    VIEWCHK(hipMemcpyAsync(out, spec_out_, (size_t)Nl_ * N_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    VIEWCHK(hipStreamSynchronize(st_));
}

void Viewer::download_image(double *out) {
    if (!has_image_) throw BpmError("no image has been computed yet (a colour matching table is required)");
    VIEWCHK(hipMemcpyAsync(out, image_, (size_t)3 * N_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    VIEWCHK(hipStreamSynchronize(st_));
}

}  // namespace nmk
