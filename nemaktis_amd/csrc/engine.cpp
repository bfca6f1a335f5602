#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <utility>
#include <vector>

#include "kernels.h"
#include "staging.h"
#include "vti.h"

namespace nmk {

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

namespace {

const double kPi = 3.1415926535897932;

template <class P>
struct DevBuf {
    P *ptr = nullptr;
    size_t count = 0;
    DeviceAllocator *source = nullptr;  // nullptr: hipMalloc
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            if (source) source->release(ptr);
            else (void)hipFree(ptr);
        }
        ptr = nullptr;
    }
    void alloc(size_t n) {
        release();
        count = 0;
        if (!n) return;
        if (source) ptr = static_cast<P *>(source->allocate(n * sizeof(P)));
        else HIPCHK(hipMalloc((void **)&ptr, n * sizeof(P)));
        count = n;
    }
    // exchanges the allocations of two buffers of the same source
    void swap(DevBuf &o) {
        std::swap(ptr, o.ptr);
        std::swap(count, o.count);
        std::swap(source, o.source);
    }
    size_t bytes() const { return count * sizeof(P); }
    size_t footprint() const { return (bytes() + kAllocAlign - 1) / kAllocAlign * kAllocAlign; }
};

template <class... B>
void bind_source(DeviceAllocator *a, B &...b) {
    ((b.source = a), ...);
}

// Catmull-Rom weights of a fractional shift nu (SimpleShiftOperator.cpp:27-30)
void cr_weights(double nu, double *w) {
    w[0] = -nu * ((1 - nu) * (1 - nu)) / 2;
    w[1] = (1 - nu) * (1 + nu - 3 * nu * nu / 2);
    w[2] = nu * (1 + 4 * nu - 3 * nu * nu) / 2;
    w[3] = -nu * nu * (1 - nu) / 2;
}

// one shift as a 7-tap filter and two successive shifts composed (w * w)
void compose_taps(const double *w, int o, double *w1, int &o1, double *w2, int &o2) {
    for (int t = 0; t < 7; t++) w1[t] = w2[t] = 0;
    for (int t = 0; t < 4; t++) w1[t] = w[t];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) w2[i + j] += w[i] * w[j];
    o1 = o;
    o2 = 2 * o;
}

// a 7-tap filter starting at offset o that returns its input unchanged
bool identity_taps(const double *w, int o) {
    for (int t = 0; t < 7; t++)
        if (w[t] != (o + t == 0 ? 1.0 : 0.0)) return false;
    return true;
}

struct Sizes {
    int nwl, nq, batch, nslots_max, nprops_max;
    int64_t Nfull, NP;
    int coef_batch, ring;  // slabs per coefficient launch, slab buffers of coefficients (2 x coef_batch)
};

Sizes sizes_of(const Problem &p, const Options &o) {
    Sizes s;
    s.nwl = (int)p.wl.size();
    s.nq = (int)p.qx.size();
    s.batch = o.wl_batch <= 0 ? s.nwl : std::min(o.wl_batch, s.nwl);
    s.nslots_max = s.batch;
    s.nprops_max = s.batch * s.nq * 2;
    s.Nfull = (int64_t)p.Nx * p.Ny;
    int Mx = p.Nx > 3 ? p.Nx - 1 : p.Nx, My = p.Ny > 3 ? p.Ny - 1 : p.Ny;
    s.NP = (int64_t)Mx * My;
    // slabs per coefficient launch (on the march stream): the launch of a small
    // slab is a fraction of one round of blocks, and several slabs per launch
    // fill the device (64^2 x 20: 1.26 -> 1.13 ms per step with 8).  But a
    // batch is computed ahead of its slabs, and the solvers depend on reading
    // planes that are still in the Infinity Cache: batches of 156 MB already
    // slowed the 512^2 FP32 step by 3.5 %, so a batch holds at most 32 MB of
    // coefficients (and at most 8 slabs, about 8192 blocks).
    const int64_t blocks = slab_coef_tiles(Mx, My) * s.nslots_max;
    const int64_t slab_bytes = (int64_t)s.nslots_max * cf_slot_size(Mx, My) * (o.precision == 32 ? 8 : 16);
    int B = 1;
    if (o.coef_batch > 0) {
        B = o.coef_batch;
    } else {
        while (B < 8 && 2 * B * blocks <= 8192 && 2 * B * slab_bytes <= (int64_t(32) << 20)) B *= 2;
    }
    s.coef_batch = std::max(1, std::min({B, std::max(1, p.Nz - 1), (int)(65535 / std::max(1, s.nslots_max))}));
    s.ring = 2 * s.coef_batch;
    return s;
}

constexpr int kNrefBlocks = 1024;

// line solves of at most 256 chunks of 16 (FP32) or 8 (FP64) points
void check_mesh_limits(const Problem &p, const Options &o) {
    const int max_n = o.precision == 32 ? 4097 : 2049;
    if (p.Nx > max_n || p.Ny > max_n)
        throw BpmError("Mesh dimensions above " + std::to_string(max_n) + " points per axis are not supported (" +
                       std::to_string(o.precision) + "-bit path)");
}

int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p *= 2;
    return p;
}

template <class T>
class RunnerT final : public Runner {
public:
    RunnerT(const Problem &p, const Options &o, hipStream_t external) : p_(p), o_(o) {
        if (o.device >= 0) HIPCHK(hipSetDevice(o.device));
        if (external) {
            st_ = external;
        } else {
            HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
            own_stream_ = true;
        }
        sz_ = sizes_of(p, o);
        bind_source(o.allocator, bulk_, bt_, lc_, weights_, q_, iso_t_, iso_n_, partials_, mask_, twx_, twy_, E_, sp_, pp_, cf_,
                    xt_, fb_[0], fb_[1], fb_[2], fb_[3], flag_);
        g_.Nx = p.Nx;
        g_.Ny = p.Ny;
        g_.Nz = p.Nz;
        g_.px = p.Nx > 3;
        g_.py = p.Ny > 3;
        g_.Mx = g_.px ? p.Nx - 1 : p.Nx;
        g_.My = g_.py ? p.Ny - 1 : p.Ny;
        g_.field_dim = p.field_dim;
        g_.Nfull = sz_.Nfull;
        g_.NP = sz_.NP;
        check_mesh_limits(p, o);
        plan_x_ = make_fft_plan(p.Nx);
        plan_y_ = make_fft_plan(p.Ny);

        // every buffer is allocated before anything is uploaded, so that a
        // run that does not fit its allocator fails before any device work
        const bool dev_in = o.dev.lc != nullptr;  // inputs and output already on the device
        if (dev_in && p.bulk) throw BpmError("bulk output needs host arrays (its planes are written to a file)");
        if (!dev_in) lc_.alloc((size_t)p.voxels() * p.field_dim);
        if (dev_in ? o.dev.mask != nullptr : !p.mask.empty()) mask_.alloc((size_t)p.voxels());
        const bool dev_weights = !dev_in && p.weights_on_device && !p.mask.empty();
        if (dev_weights) {
            weights_.alloc(2 * (size_t)sz_.Nfull * (size_t)(p.Nz - 1));
            flag_.alloc(1);
        } else if (!dev_in && !p.weights.empty()) {
            weights_.alloc(p.weights.size());
        }
        q_.alloc(2 * (size_t)sz_.nq);
        iso_t_.alloc(std::max<size_t>(1, p.iso_thick.size()));
        iso_n_.alloc(std::max<size_t>(1, p.iso_index.size()));
        twx_.alloc(p.Nx);
        twy_.alloc(p.Ny);
        sp_.alloc(sz_.nslots_max);
        pp_.alloc(sz_.nprops_max);
        partials_.alloc((size_t)kNrefBlocks * sz_.nslots_max);
        cf_.alloc((size_t)sz_.ring * sz_.nslots_max * cf_slot_size(g_.Mx, g_.My));
        set_tuning(o.tuning.x_chunks > 0 ? o.tuning : default_tuning(o.precision, g_.Mx, g_.My, sz_.nprops_max));
        for (auto &b : fb_) b.alloc((size_t)sz_.nprops_max * 2 * sz_.NP);
        if (!dev_in) E_.alloc((size_t)sz_.nwl * sz_.nq * 4 * sz_.Nfull);
        if (p.bulk) {
            // every plane of every propagation stays on the device until the
            // run writes its file (the reference holds them in host memory)
            bulk_.alloc((size_t)sz_.nwl * sz_.nq * 2 * p.Nz * 2 * sz_.Nfull);
            bt_.alloc((size_t)sz_.nprops_max * 2 * sz_.NP);
        }
        lc_p_ = dev_in ? o.dev.lc : lc_.ptr;
        mask_p_ = mask_.ptr;
        w_p_ = dev_in ? o.dev.weights : weights_.ptr;
        E_p_ = dev_in ? o.dev.E : E_.ptr;
        if (!o.profile && !o.no_side_stream) HIPCHK(hipStreamCreateWithFlags(&st2_, hipStreamNonBlocking));

        // inputs resident in HBM
        if (!dev_in) staged_h2d(lc_.ptr, p.lc, lc_.bytes(), st_);
        if (mask_.ptr) {
            if (dev_in) launch_mask_flags(o.dev.mask, mask_.ptr, p.voxels(), st_);
            else staged_h2d(mask_.ptr, p.mask.data(), mask_.bytes(), st_);
        }
        if (dev_weights) {
            // the interface weights of the mask formula, from the LC flags just uploaded
            HIPCHK(hipMemsetAsync(flag_.ptr, 0, sizeof(int), st_));
            launch_interface_weights(p.weight_prog, mask_.ptr, p.Nx, p.Ny, p.Nz, p.dx, p.dy, p.dz, weights_.ptr,
                                     flag_.ptr, st_);
            HIPCHK(hipMemcpyAsync(&mismatch_, flag_.ptr, sizeof(int), hipMemcpyDeviceToHost, st_));
        } else if (weights_.ptr) {
            staged_h2d(weights_.ptr, p.weights.data(), weights_.bytes(), st_);
        }
        hq_.clear();
        for (int i = 0; i < sz_.nq; i++) {
            hq_.push_back(p.qx[i]);
            hq_.push_back(p.qy[i]);
        }
        HIPCHK(hipMemcpyAsync(q_.ptr, hq_.data(), q_.bytes(), hipMemcpyHostToDevice, st_));
        if (!p.iso_thick.empty()) {
            HIPCHK(hipMemcpyAsync(iso_t_.ptr, p.iso_thick.data(), p.iso_thick.size() * 8, hipMemcpyHostToDevice, st_));
            HIPCHK(hipMemcpyAsync(iso_n_.ptr, p.iso_index.data(), p.iso_index.size() * 8, hipMemcpyHostToDevice, st_));
        }
        auto twiddles = [&](int n, DevBuf<c64> &d, std::vector<c64> &h) {
            h.resize(n);
            for (int k = 0; k < n; k++) {
                // exp(-2 pi i k / n), argument reduced to [0, pi/4] octants for accuracy
                long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
                h[k] = mk((double)std::cos(a), (double)std::sin(a));
            }
            HIPCHK(hipMemcpyAsync(d.ptr, h.data(), d.bytes(), hipMemcpyHostToDevice, st_));
        };
        twiddles(p.Nx, twx_, htwx_);
        twiddles(p.Ny, twy_, htwy_);
        HIPCHK(hipStreamSynchronize(st_));
        if (dev_weights && mismatch_) {
            // the reference prints the notice and keeps every weight at 0.5
            w_p_ = nullptr;
            notice_ = kMaskMismatchNotice;
        }
    }

    ~RunnerT() override {
        // work queued by a propagation that threw (or by a caller's stream)
        // may still read or write the buffers released below: finish it first
        (void)hipStreamSynchronize(st_);
        if (st2_) (void)hipStreamSynchronize(st2_);
        drop_graphs();
        for (auto &e : evs_) (void)hipEventDestroy(e);
        for (auto &pr : prof_pool_) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        for (auto &x : sync_evs_) (void)hipEventDestroy(x);
        if (st2_) (void)hipStreamDestroy(st2_);
        if (own_stream_) (void)hipStreamDestroy(st_);
    }

    hipStream_t stream() const override { return st_; }
    const std::string &notice() const override { return notice_; }
    const c64 *device_output() const override { return E_p_; }

    int64_t footprint() const override {
        size_t b = lc_.footprint() + weights_.footprint() + q_.footprint() + iso_t_.footprint() + iso_n_.footprint() +
                   partials_.footprint() + mask_.footprint() + twx_.footprint() + twy_.footprint() +
                   E_.footprint() + sp_.footprint() + pp_.footprint() + cf_.footprint() + xt_.footprint() +
                   bulk_.footprint() + bt_.footprint() + flag_.footprint();
        for (auto &f : fb_) b += f.footprint();
        return (int64_t)b;
    }

    void upload_fields(const c64 *E_host) override {
        staged_h2d(E_p_, E_host, out_bytes(), st_);
        HIPCHK(hipStreamSynchronize(st_));
    }

    void download(c64 *E_host) override { staged_d2h(E_host, E_p_, out_bytes(), st_); }

    void propagate(Stats *stats) override {
        // no_timing: no events at all (the run creates no HIP objects); the
        // stage times stay 0 and the stream is synchronised instead
        const bool timing = !o_.no_timing;
        auto mark = [&](size_t i) {
            if (!timing) return (hipEvent_t) nullptr;
            hipEvent_t e = ev(i);
            HIPCHK(hipEventRecord(e, st_));
            return e;
        };
        auto elapsed = [&](hipEvent_t a, hipEvent_t b) {
            float t = 0;
            if (a && b) HIPCHK(hipEventElapsedTime(&t, a, b));
            return (double)t;
        };
        hipEvent_t e0 = mark(0);
        double march_ms = 0, setup_ms = 0, out_ms = 0;
        if (stats) *stats = Stats();
        stats_ = stats;
        for (int w0 = 0; w0 < sz_.nwl; w0 += sz_.batch) {
            const int nslots = std::min(sz_.batch, sz_.nwl - w0);
            const int nprops = nslots * sz_.nq * 2;
            hipEvent_t a = mark(2);
            setup_batch(w0, nslots, nprops);
            hipEvent_t b = mark(3);
            C2<T> *state = march(nslots, nprops);
            hipEvent_t c = mark(4);
            output_batch(w0, nslots, nprops, state);
            hipEvent_t d = mark(5);
            if (d) HIPCHK(hipEventSynchronize(d));
            else HIPCHK(hipStreamSynchronize(st_));
            resolve_profile();
            setup_ms += elapsed(a, b);
            march_ms += elapsed(b, c);
            out_ms += elapsed(c, d);
        }
        hipEvent_t e1 = mark(1);
        if (e1) HIPCHK(hipEventSynchronize(e1));
        if (p_.bulk) write_bulk();
        if (stats) {
            stats->total_ms = elapsed(e0, e1);
            stats->setup_ms = setup_ms;
            stats->march_ms = march_ms;
            stats->output_ms = out_ms;
            stats->updates = (double)p_.Nx * p_.Ny * (p_.Nz - 1) * (double)sz_.nwl * sz_.nq * 2;
            stats->footprint_bytes = footprint();
        }
        stats_ = nullptr;
    }

private:
    // the bulk planes to <Results folder name>/<Base name>.vti
    // (PythonBackend/bpm_backend.cpp:44-47, PostprocessorLib/VolumeOutput.cpp:19-32)
    void write_bulk() {
        BulkGeometry bg;
        bg.Nx = p_.Nx;
        bg.Ny = p_.Ny;
        bg.Nz = p_.Nz;
        bg.dx = p_.dx;
        bg.dy = p_.dy;
        bg.dz = p_.dz;
        bg.wl = p_.wl;
        bg.qx = p_.qx;
        bg.qy = p_.qy;
        const size_t blk = (size_t)p_.Nz * 2 * sz_.Nfull;
        write_bulk_vti(p_.bulk_path, p_.bulk_dir, bg, [&](size_t k, c64 *dst) {
            staged_d2h(dst, bulk_.ptr + k * blk, blk * sizeof(c64), st_);
        });
    }

    Problem p_;
    Options o_;
    Sizes sz_;
    Geo g_;
    hipStream_t st_ = nullptr, st2_ = nullptr;  // march stream, coefficient stream
    bool own_stream_ = false;
    std::vector<hipEvent_t> sync_evs_;
    FftPlan plan_x_, plan_y_;
    DevBuf<double> lc_, weights_, q_, iso_t_, iso_n_, partials_;
    DevBuf<uint8_t> mask_;
    DevBuf<int> flag_;      // mismatch flag of the device interface weights
    int mismatch_ = 0;
    std::string notice_;
    DevBuf<c64> twx_, twy_, E_;
    std::vector<c64> htwx_, htwy_;
    DevBuf<SlotParams> sp_;
    DevBuf<PropParams> pp_;
    DevBuf<C2<T>> cf_, xt_;
    DevBuf<C2<T>> fb_[4];
    DevBuf<c64> bulk_;   // bulk output planes (p.bulk)
    // inputs and output as the kernels see them: the runner's own buffers, or
    // the caller's device arrays (Options::dev)
    const double *lc_p_ = nullptr, *w_p_ = nullptr;
    const uint8_t *mask_p_ = nullptr;
    c64 *E_p_ = nullptr;
    std::vector<double> hq_;
    size_t out_bytes() const { return (size_t)sz_.nwl * sz_.nq * 4 * sz_.Nfull * sizeof(c64); }
    DevBuf<C2<T>> bt_;   // one boundary-shifted plane of the bulk output
    LinePlan lpx_{}, lpy_{};
    Tuning tune_{};
    std::vector<SlotParams> hsp_;
    std::vector<PropParams> hpp_;
    bool unshifted_[3] = {false, false, false};  // by number of shifts of a boundary
    std::vector<hipEvent_t> evs_;
    struct Captured {
        hipGraphExec_t exec;
        hipGraph_t graph;
        C2<T> *state;  // buffer holding the final state after a replay
    };
    std::map<std::pair<int, int>, Captured> graphs_;
    Stats *stats_ = nullptr;

    hipEvent_t ev(size_t i) {
        while (evs_.size() <= i) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evs_.push_back(e);
        }
        return evs_[i];
    }

    // kernel-family timing in profile mode: an event pair around every launch,
    // resolved once the propagation has finished (no host sync in between)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pool_;
    std::vector<int> prof_fam_;
    template <class F>
    void timed(int fam, F &&launch) {
        if (!o_.profile || !stats_) {
            launch();
            return;
        }
        const size_t k = prof_fam_.size();
        while (prof_pool_.size() <= k) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            prof_pool_.emplace_back(a, b);
        }
        HIPCHK(hipEventRecord(prof_pool_[k].first, st_));
        launch();
        HIPCHK(hipEventRecord(prof_pool_[k].second, st_));
        prof_fam_.push_back(fam);
    }
    void resolve_profile() {
        if (!stats_) return;
        for (size_t k = 0; k < prof_fam_.size(); k++) {
            HIPCHK(hipEventSynchronize(prof_pool_[k].second));
            float t;
            HIPCHK(hipEventElapsedTime(&t, prof_pool_[k].first, prof_pool_[k].second));
            stats_->family_ms[prof_fam_[k]] += t;
            stats_->family_launches[prof_fam_[k]] += 1;
        }
        prof_fam_.clear();
    }

    void setup_batch(int w0, int nslots, int nprops) {
        const Problem &p = p_;
        hsp_.assign(nslots, SlotParams());
        for (int s = 0; s < nslots; s++) {
            const int w = w0 + s;
            SlotParams &sp = hsp_[s];
            std::memset(&sp, 0, sizeof(sp));
            sp.k0 = 2 * kPi / p.wl[w];
            sp.dX = sp.k0 * p.dx;
            sp.dY = sp.k0 * p.dy;
            sp.dZ = sp.k0 * p.dz;
            sp.no = p.no[w];
            sp.ne = p.ne[w];
            sp.ne_imag = p.ne_imag[w];
            sp.nhost = p.nhost[w];
            sp.eperp = sp.no * sp.no;
            sp.ea = sp.ne * sp.ne - sp.eperp;
            sp.ehost = sp.nhost * sp.nhost;
            sp.alpha_im = p.ne_imag[w] / (p.ne[w] - p.no[w]);  // NaN/inf when ne == no, as the reference
            sp.nin = p.nin[w];
            sp.nout_fresnel = p.nout_fresnel[w];
            sp.nout_filter = p.nout[w];
            sp.iwl = w;
        }
        timed(KF_SETUP, [&] {
            HIPCHK(hipMemcpyAsync(sp_.ptr, hsp_.data(), nslots * sizeof(SlotParams), hipMemcpyHostToDevice, st_));
            launch_nref(lc_p_, mask_p_, g_, sp_.ptr, nslots, partials_.ptr, kNrefBlocks, st_);
            HIPCHK(hipMemcpyAsync(hsp_.data(), sp_.ptr, nslots * sizeof(SlotParams), hipMemcpyDeviceToHost, st_));
            HIPCHK(hipStreamSynchronize(st_));
        });
        // per-propagation shift parameters (SimpleShiftOperator.cpp:3-50) and
        // the product of all shift phase factors, in FP64
        hpp_.assign(nprops, PropParams());
        for (int s = 0; s < nslots; s++)
            for (int iq = 0; iq < sz_.nq; iq++)
                for (int pol = 0; pol < 2; pol++) {
                    PropParams &pp = hpp_[pol + 2 * (iq + sz_.nq * s)];
                    std::memset(&pp, 0, sizeof(pp));
                    const SlotParams &sp = hsp_[s];
                    const double nref = sp.nref, qx = p.qx[iq], qy = p.qy[iq];
                    pp.slot = s;
                    pp.iq = iq;
                    pp.pol = pol;
                    pp.qx = qx;
                    pp.qy = qy;
                    const double th = 0.5 * kPi * pol;
                    pp.ex = std::cos(th);
                    pp.ey = std::sin(th);
                    if (p.Nx > 3) {
                        double xs = -sp.dZ * qx / (2 * nref * sp.dX);
                        double ms = std::floor(xs);
                        pp.ox = (int)ms - 1;
                        cr_weights(xs - ms, pp.wx);
                    } else {
                        pp.wx[1] = 1;
                    }
                    if (p.Ny > 3) {
                        double ys = -sp.dZ * qy / (2 * nref * sp.dX);  // deltaX, as the reference (:34)
                        double ms = std::floor(ys);
                        pp.oy = (int)ms - 1;
                        cr_weights(ys - ms, pp.wy);
                    } else {
                        pp.wy[1] = 1;
                    }
                    compose_taps(pp.wx, pp.ox, pp.w1x, pp.o1x, pp.w2x, pp.o2x);
                    compose_taps(pp.wy, pp.oy, pp.w1y, pp.o1y, pp.w2y, pp.o2y);
                    pp.ph_half = -sp.dZ * (qx * qx + qy * qy) / (4 * nref);
                    const double phi = -sp.dZ * (qx * qx + qy * qy) / (4 * nref) * 2.0 * (p.Nz - 1);
                    pp.ph_re = std::cos(phi);
                    pp.ph_im = std::sin(phi);
                }
        // launches whose propagations are all unshifted (normal incidence) take the pointwise boundary
        for (int ns = 1; ns <= 2; ns++) {
            unshifted_[ns] = true;
            for (const PropParams &pp : hpp_) {
                const bool ix = !g_.px || identity_taps(ns == 2 ? pp.w2x : pp.w1x, ns == 2 ? pp.o2x : pp.o1x);
                const bool iy = !g_.py || identity_taps(ns == 2 ? pp.w2y : pp.w1y, ns == 2 ? pp.o2y : pp.o1y);
                unshifted_[ns] = unshifted_[ns] && ix && iy;
            }
        }
        timed(KF_SETUP, [&] {
            HIPCHK(hipMemcpyAsync(pp_.ptr, hpp_.data(), nprops * sizeof(PropParams), hipMemcpyHostToDevice, st_));
            C2<T> *F = fb_[0].ptr;
            if (p.beam == BeamKind::None) {
                const c64 *Ein = E_p_ + (size_t)w0 * sz_.nq * 4 * sz_.Nfull;
                launch_pack_fields<T>(Ein, g_, pp_.ptr, nprops, sz_.nq, F, st_);
            } else {
                launch_init_fields<T>(lc_p_, mask_p_, g_, sp_.ptr, pp_.ptr, nprops,
                                      p.beam == BeamKind::Gaussian ? 1 : 0, p.waist, p.dx, p.dy, F, st_);
            }
        });
    }

    // the z-march over all slabs; returns the buffer holding the final state
    C2<T> *march(int nslots, int nprops) {
        if (o_.use_graph && !o_.profile && p_.Nz > 1) {
            auto key = std::make_pair(nslots, nprops);
            auto it = graphs_.find(key);
            if (it == graphs_.end()) {
                hipGraph_t graph;
                HIPCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
                C2<T> *state = nullptr;
                try {
                    state = march_launches(nslots, nprops);
                } catch (...) {
                    hipGraph_t junk;
                    (void)hipStreamEndCapture(st_, &junk);
                    if (junk) (void)hipGraphDestroy(junk);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(st_, &graph));
                hipGraphExec_t exec;
                HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                it = graphs_.emplace(key, Captured{exec, graph, state}).first;
            }
            HIPCHK(hipGraphLaunch(it->second.exec, st_));
            return it->second.state;
        }
        return march_launches(nslots, nprops);
    }

    hipEvent_t sync_ev(size_t i) {
        while (sync_evs_.size() <= i) {
            hipEvent_t x;
            HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            sync_evs_.push_back(x);
        }
        return sync_evs_[i];
    }

    // The z-march.  Coefficient planes live in a ring of sz_.ring slab
    // buffers.  On the march stream (coef_serial, the built-in choice) each
    // batch of sz_.coef_batch slabs is computed in one launch right before the
    // boundary that first reads it, so that its slabs march while their planes
    // are still in the Infinity Cache (512^2 FP64: 155 MB per slab, 26.1 ->
    // 24.8 ms per step against computing them a slab ahead).  On a second
    // stream (an autotuner option) slab k+2's are computed while slab k
    // marches, and the boundary joins the two streams (about 10 us per join).
    C2<T> *march_launches(int nslots, int nprops) {
        const int nsl = p_.Nz - 1, nw = p_.n_woodbury, R = sz_.ring, CB = sz_.coef_batch;
        const size_t cfsz = (size_t)sz_.nslots_max * cf_slot_size(g_.Mx, g_.My);
        const size_t xtsz = xt_.count / R;
        auto CF = [&](int k) { return cf_.ptr + (size_t)(k % R) * cfsz; };
        auto XT = [&](int k) { return xt_.ptr ? xt_.ptr + (size_t)(k % R) * xtsz : nullptr; };
        const bool overlap = st2_ != nullptr && !tune_.coef_serial;
        hipStream_t cs = overlap ? st2_ : st_;
        size_t evn = 0;
        auto record = [&](hipStream_t s) {
            hipEvent_t x = sync_ev(evn++);
            HIPCHK(hipEventRecord(x, s));
            return x;
        };
        // the event that slab k's coefficients are ready (only when they run on
        // the second stream; on the march stream the order is the stream's)
        // coefficients of slabs k0 .. k0+n-1 in one launch on stream s
        auto coefs = [&](int k0, int n, hipStream_t s) {
            timed(KF_COEFS, [&] {
                launch_slab_coefs<T>(lc_p_, mask_p_, w_p_, g_, sp_.ptr, nslots, k0, n, cf_.ptr, (int64_t)cfsz,
                                     xt_.ptr, (int64_t)xtsz, R, lpx_, s);
            });
        };
        auto coef = [&](int k) {
            coefs(k, 1, cs);
            return overlap ? record(cs) : (hipEvent_t) nullptr;
        };
        if (nsl == 0) {
            // a single plane: no slab to march, the entrance fields go to the
            // output as they are (BPMIteration.cpp:127 runs its loop Nz-1 times)
            if (bulk_.ptr) {
                const int beam = p_.beam == BeamKind::None ? 2 : p_.beam == BeamKind::Gaussian ? 1 : 0;
                launch_bulk_entrance(lc_p_, mask_p_, g_, sp_.ptr, pp_.ptr, nprops, beam, p_.waist, p_.dx, p_.dy,
                                     E_p_, sz_.nq, bulk_.ptr, st_);
            }
            return fb_[0].ptr;
        }
        if (overlap) HIPCHK(hipStreamWaitEvent(cs, record(st_), 0));
        std::vector<hipEvent_t> ready(nsl, nullptr);
        if (overlap) {
            ready[0] = coef(0);
            if (nsl > 1) ready[1] = coef(1);
        } else {
            coefs(0, std::min(CB, nsl), st_);
        }

        C2<T> *cur = fb_[1].ptr, *other = fb_[0].ptr, *A = fb_[2].ptr, *B = fb_[3].ptr;
        // XCD work order (measured): a lone polarisation pair shares L2 best
        // across the two components; batches whose field planes total up to
        // about 32 MiB do best with the pairs innermost (the wavevectors of a
        // wavelength read the same coefficient rows on one XCD), larger ones
        // with the pairs outermost (each XCD keeps to its own pairs' fields)
        // in FP32, and in FP64 with each wavelength's pairs grouped (bit 3:
        // a third of the HBM fetches, which only the FP64 solvers turn into
        // time)
        const size_t plane_bytes = (size_t)nprops * g_.Nx * g_.Ny * sizeof(C2<T>);
        const int ord0 = tune_.order >= 0 ? tune_.order
                         : nprops <= 2 ? 3
                         : plane_bytes <= (size_t(32) << 20) ? 6
                         : sizeof(T) == 8 ? 10 : 0;
        const int ord = ord0 | (sz_.nq << 4);  // group size of order bit 3: the pairs of one slot
        if (ready[0]) HIPCHK(hipStreamWaitEvent(st_, ready[0], 0));
        // bulk plane 0: the entrance fields after the input interface
        if (bulk_.ptr) {
            const int beam = p_.beam == BeamKind::None ? 2 : p_.beam == BeamKind::Gaussian ? 1 : 0;
            launch_bulk_entrance(lc_p_, mask_p_, g_, sp_.ptr, pp_.ptr, nprops, beam, p_.waist, p_.dx, p_.dy,
                                 E_p_, sz_.nq, bulk_.ptr, st_);
        }
        timed(KF_BOUNDARY, [&] { launch_boundary<T>(other, cur, nullptr, CF(0), g_, pp_.ptr, nprops, 1, unshifted_[1], st_); });
        for (int k = 0; k < nsl; k++) {
            // x half-steps: other = (I - mu A_x)^-1 (I + mu A_y) cur by its Neumann series
            // y0 + y1 + ... + y_nw: y0 goes to A, the first Neumann launch adds it to y1
            // into other, the last one writes only the sum (in place)
            if (g_.px) {
                timed(KF_SOLVE_X, [&] {
                    launch_xsolve<T>(0, cur, nw > 0 ? A : other, nullptr, nullptr, CF(k), XT(k), g_, pp_.ptr, nprops, lpx_, tune_.x_rows, ord, tune_.lean, tune_.x_both & 1, st_);
                });
                C2<T> *a = A, *b = B;
                for (int it = 0; it < nw; it++) {
                    timed(KF_SOLVE_X, [&] {
                        launch_xsolve<T>(1, a, it + 1 < nw ? b : nullptr, it == 0 ? a : other, other, CF(k), XT(k), g_, pp_.ptr, nprops, lpx_, tune_.x_rows, ord, tune_.lean, (tune_.x_both >> 1) & 1, st_);
                    });
                    std::swap(a, b);
                }
            } else {
                timed(KF_EXPLICIT, [&] { launch_explicit<T>(1, cur, other, CF(k), g_, pp_.ptr, nprops, st_); });
            }
            // y half-steps: cur = (I - mu A_y)^-1 (I + mu A_x) other by its Neumann series
            if (g_.py) {
                timed(KF_SOLVE_Y, [&] {
                    launch_ysolve<T>(0, other, nw > 0 ? A : cur, nullptr, nullptr, CF(k), g_, pp_.ptr, nprops, lpy_, tune_.y_cols, ord, tune_.lean, tune_.y_pair, st_);
                });
                C2<T> *a = A, *b = B;
                for (int it = 0; it < nw; it++) {
                    timed(KF_SOLVE_Y, [&] {
                        launch_ysolve<T>(1, a, it + 1 < nw ? b : nullptr, it == 0 ? a : cur, cur, CF(k), g_, pp_.ptr, nprops, lpy_, tune_.y_cols, ord, tune_.lean, tune_.y_pair, st_);
                    });
                    std::swap(a, b);
                }
            } else {
                timed(KF_EXPLICIT, [&] { launch_explicit<T>(0, other, cur, CF(k), g_, pp_.ptr, nprops, st_); });
            }
            // bulk plane k+1: the step's closing P_k, Sh applied to a copy
            if (bulk_.ptr) {
                launch_boundary<T>(cur, bt_.ptr, CF(k), nullptr, g_, pp_.ptr, nprops, 1, unshifted_[1], st_);
                launch_bulk_plane<T>(bt_.ptr, g_, sp_.ptr, pp_.ptr, nprops, k + 1, sz_.nq, bulk_.ptr, st_);
            }
            // P_k, Sh | Sh, P_{k+1}
            if (k + 1 < nsl) {
                // coefficients on the march stream: the batch that starts at slab k+1 is
                // computed here, just before the boundary and the slabs that read it (so
                // it is still in the Infinity Cache), into the buffers of the batch before
                // the current one, whose last reader has run
                if (!overlap && (k + 1) % CB == 0) coefs(k + 1, std::min(CB, nsl - k - 1), st_);
                if (ready[k + 1]) HIPCHK(hipStreamWaitEvent(st_, ready[k + 1], 0));
                timed(KF_BOUNDARY, [&] { launch_boundary<T>(cur, other, CF(k), CF(k + 1), g_, pp_.ptr, nprops, 2, unshifted_[2], st_); });
            } else {
                timed(KF_BOUNDARY, [&] { launch_boundary<T>(cur, other, CF(k), nullptr, g_, pp_.ptr, nprops, 1, unshifted_[1], st_); });
            }
            std::swap(cur, other);
            if (overlap && k + 2 < nsl) {
                // CF(k) and XT(k) are free once the boundary above has run
                HIPCHK(hipStreamWaitEvent(cs, record(st_), 0));
                ready[k + 2] = coef(k + 2);
            }
        }
        if (overlap) HIPCHK(hipStreamWaitEvent(st_, record(cs), 0));
        return cur;
    }

    void drop_graphs() {
        for (auto &kv : graphs_) {
            (void)hipGraphExecDestroy(kv.second.exec);
            (void)hipGraphDestroy(kv.second.graph);
        }
        graphs_.clear();
    }

    // validates t, then switches to it (the runner keeps its state when t is
    // refused).  The captured graphs hold the buffer pointers of the old
    // shapes, so they are destroyed before any buffer changes; a larger x-line
    // row buffer is allocated beside the old one and swapped in only once the
    // allocation has succeeded.
    void set_tuning(const Tuning &t) {
        const LinePlan lx = g_.px ? make_line_plan(g_.Mx, t.x_chunks) : lpx_;
        const LinePlan ly = g_.py ? make_line_plan(g_.My, t.y_chunks) : lpy_;
        if (g_.px && ((t.x_both ? 4 : 2) * lx.P * t.x_rows > 512 ||
                      xsolve_lds_bytes<T>(lx, t.x_rows, t.x_both ? 2 : 1) > max_block_lds()))
            throw std::runtime_error("x line solver tuning exceeds the block limits");
        if (g_.py && ((t.y_pair ? 2 : 1) * ly.P * t.y_cols > 512 ||
                      ysolve_lds_bytes<T>(ly, t.y_cols, t.y_pair) > max_block_lds()))
            throw std::runtime_error("y line solver tuning exceeds the block limits");
        drop_graphs();
        if (g_.px) {
            const size_t need = (size_t)sz_.ring * sz_.nslots_max * 2 * 3 * g_.My * lx.lmax * lx.P;
            if (need > xt_.count) {
                DevBuf<C2<T>> grown;
                grown.source = xt_.source;
                grown.alloc(need);  // throws with xt_ untouched
                xt_.swap(grown);    // the old rows are released with `grown`
            }
        }
        tune_ = t;
        lpx_ = lx;
        lpy_ = ly;
    }

    void retune(const Tuning &t) override {
        HIPCHK(hipStreamSynchronize(st_));
        if (st2_) HIPCHK(hipStreamSynchronize(st2_));
        set_tuning(t.x_chunks > 0 ? t : default_tuning(o_.precision, g_.Mx, g_.My, sz_.nprops_max));
    }

    void output_batch(int w0, int nslots, int nprops, const C2<T> *state) {
        c64 *Eb = E_p_ + (size_t)w0 * sz_.nq * 4 * sz_.Nfull;
        timed(KF_OUTPUT, [&] {
            if (p_.Nz > 1) {
                launch_output<T>(state, lc_p_, mask_p_, g_, sp_.ptr, pp_.ptr, nprops, sz_.nq, Eb, st_);
            } else {
                const int beam = p_.beam == BeamKind::None ? 2 : p_.beam == BeamKind::Gaussian ? 1 : 0;
                launch_output_entrance(lc_p_, mask_p_, g_, sp_.ptr, pp_.ptr, nprops, beam, p_.waist, p_.dx, p_.dy, Eb,
                                       st_);
            }
        });
        timed(KF_FILTER, [&] {
            const int planes = nslots * sz_.nq * 4;
            launch_fft2(Eb, p_.Nx, p_.Ny, planes, -1, twx_.ptr, twy_.ptr, plan_x_, plan_y_, st_);
            launch_filter(Eb, g_, sp_.ptr, nslots, q_.ptr, sz_.nq, iso_t_.ptr, iso_n_.ptr, (int)p_.iso_thick.size(),
                          p_.dx, p_.dy, p_.z_foc, p_.NA, st_);
            launch_fft2(Eb, p_.Nx, p_.Ny, planes, +1, twx_.ptr, twy_.ptr, plan_x_, plan_y_, st_);
        });
    }
};

}  // namespace

// Built-in line-solver shapes.  The chunks per line trade parallelism
// (about 128 K chunk threads fill the GPU) against the length of the
// per-chunk chains (at most 16 points, 8 in FP64, whose longer chains spill
// registers): what the autotuner picks at 512^2 and 1024^2 with two
// propagations (64 chunks) and on the 294-propagation 256^2 sweep (16).
// y strips are as wide as the block allows up to one coefficient tile.
Tuning default_tuning(int precision, int Mx, int My, int nprops) {
    Tuning t;
    // a batch of at least 131072 lines per launch fills the device without
    // splitting lines finely: 8-point chunks (at most 64 per line, so the
    // reduced systems stay in one wave) and the register-lean solvers (more
    // blocks in flight); measured on the 256^2 x 294-propagation FP32 sweep,
    // 270.4 -> 261.5 ms per graph replay
    const bool big = std::max(Mx, My) <= 512 && (int64_t)std::max(Mx, My) * 2 * std::max(2, nprops) >= 131072;
    const int lcap = precision == 32 && !big ? 16 : 8;
    if (big && precision == 32) t.lean = 1;
    auto chunks = [&](int M, int other) {
        const int64_t lines = (int64_t)std::max(1, other) * 2 * std::max(2, nprops);
        int par = 1;
        while (par < 64 && (int64_t)par * lines < 131072) par *= 2;
        return std::min(256, std::max(16, std::max(pow2_at_least((M + lcap - 1) / lcap), par)));
    };
    const size_t budget = max_block_lds();
    auto xlds = [&](int chunks, int rows, int nc) {
        const LinePlan lp = make_line_plan(Mx, chunks);
        return precision == 32 ? xsolve_lds_bytes<float>(lp, rows, nc) : xsolve_lds_bytes<double>(lp, rows, nc);
    };
    auto ylds = [&](int chunks, int cols) {
        const LinePlan lp = make_line_plan(My, chunks);
        return precision == 32 ? ysolve_lds_bytes<float>(lp, cols, 0) : ysolve_lds_bytes<double>(lp, cols, 0);
    };
    t.x_chunks = chunks(Mx, My);
    // both field components per block in every x launch (their right-hand
    // sides have 6 and 14 taps in the first launch, 11 and 3 in the Neumann
    // launches; blocks holding both carry equal work)
    t.x_rows = 1;
    t.x_both = 3;
    if (4 * t.x_chunks > 512 || xlds(t.x_chunks, 1, 2) > budget) {
        t.x_both = 0;
        t.x_rows = 2;
        if (Mx > 3)
            while (t.x_rows > 1 && (2 * t.x_chunks * t.x_rows > 512 || xlds(t.x_chunks, t.x_rows, 1) > budget))
                t.x_rows /= 2;
    }
    t.y_chunks = chunks(My, Mx);
    // strips at most one coefficient tile wide: a wave then reads whole tiles
    t.y_cols = std::max(1, std::min(kCfTw, 512 / t.y_chunks));
    if (My > 3)
        while (t.y_cols > 1 && (t.y_chunks * t.y_cols > 512 || ylds(t.y_chunks, t.y_cols) > budget)) t.y_cols /= 2;
    // slab coefficients on the march stream: as fast or faster than the second
    // stream on every measured batch (up to 12% on small ones, in graph replay)
    t.coef_serial = 1;
    return t;
}

int64_t estimate_footprint(const Problem &p, const Options &o) {
    // mirrors the allocations of RunnerT, buffer by buffer
    check_mesh_limits(p, o);
    const Sizes s = sizes_of(p, o);
    const int64_t fsz = o.precision == 32 ? 8 : 16;
    const int Mx = p.Nx > 3 ? p.Nx - 1 : p.Nx, My = p.Ny > 3 ? p.Ny - 1 : p.Ny;
    int64_t b = 0;
    auto add = [&](int64_t bytes) { b += (bytes + (int64_t)kAllocAlign - 1) / (int64_t)kAllocAlign * (int64_t)kAllocAlign; };
    const bool dev_in = o.dev.lc != nullptr;  // the caller's device arrays: no director, weights or output copies
    if (!dev_in) add(p.voxels() * p.field_dim * 8);
    // a problem whose mask preparation is deferred holds its mask and (with a
    // formula) the interface weights once prepared: counted at full size
    const int64_t nw = 2 * (int64_t)p.Nx * p.Ny * std::max(0, p.Nz - 1);
    if (dev_in) {
        if (o.dev.mask) add(p.voxels());
    } else {
        add(p.mask_src ? p.voxels() : (int64_t)p.mask.size());
        const bool formula = p.mask_src ? !p.mask_formula.empty() : p.weights_on_device && !p.mask.empty();
        add((formula ? nw : (int64_t)p.weights.size()) * 8);
        if (formula) add(sizeof(int));  // mismatch flag of the device weights
    }
    add((int64_t)s.nq * 2 * 8);
    add((int64_t)std::max<size_t>(1, p.iso_thick.size()) * 8);
    add((int64_t)std::max<size_t>(1, p.iso_index.size()) * 8);
    add((int64_t)p.Nx * 16);
    add((int64_t)p.Ny * 16);
    add((int64_t)s.nslots_max * (int64_t)sizeof(SlotParams));
    add((int64_t)s.nprops_max * (int64_t)sizeof(PropParams));
    add((int64_t)kNrefBlocks * s.nslots_max * 8);
    add((int64_t)s.ring * s.nslots_max * cf_slot_size(Mx, My) * fsz);
    if (p.Nx > 3) {
        const Tuning t = o.tuning.x_chunks > 0 ? o.tuning : default_tuning(o.precision, Mx, My, s.nprops_max);
        const LinePlan lp = make_line_plan(Mx, t.x_chunks);
        add((int64_t)s.ring * s.nslots_max * 2 * 3 * My * lp.lmax * lp.P * fsz);
    }
    for (int i = 0; i < 4; i++) add((int64_t)s.nprops_max * 2 * s.NP * fsz);
    if (!dev_in) add((int64_t)s.nwl * s.nq * 4 * s.Nfull * 16);
    if (p.bulk) {
        add((int64_t)s.nwl * s.nq * 2 * p.Nz * 2 * s.Nfull * 16);
        add((int64_t)s.nprops_max * 2 * s.NP * fsz);
    }
    return b;
}

void device_interface_weights(const Problem &p, std::vector<double> &w, std::string *warning) {
    if (!p.weights_on_device || p.mask.size() != (size_t)p.voxels())
        throw BpmError("device interface weights need a compiled mask formula and the LC flags");
    const size_t nw = 2 * (size_t)p.Nx * p.Ny * (size_t)std::max(0, p.Nz - 1);
    w.assign(nw, 0.5);
    if (!nw) return;
    DevBuf<uint8_t> mask;
    DevBuf<double> dw;
    DevBuf<int> flag;
    mask.alloc(p.mask.size());
    dw.alloc(nw);
    flag.alloc(1);
    int mismatch = 0;
    HIPCHK(hipMemcpy(mask.ptr, p.mask.data(), mask.bytes(), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(flag.ptr, 0, sizeof(int)));
    launch_interface_weights(p.weight_prog, mask.ptr, p.Nx, p.Ny, p.Nz, p.dx, p.dy, p.dz, dw.ptr, flag.ptr, nullptr);
    HIPCHK(hipMemcpy(w.data(), dw.ptr, dw.bytes(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&mismatch, flag.ptr, sizeof(int), hipMemcpyDeviceToHost));
    if (mismatch) {
        if (warning) *warning = kMaskMismatchNotice;
        w.clear();
    }
}

std::unique_ptr<Runner> make_runner(const Problem &p, const Options &o, hipStream_t stream) {
    if (o.precision == 32) return std::unique_ptr<Runner>(new RunnerT<float>(p, o, stream));
    if (o.precision == 64) return std::unique_ptr<Runner>(new RunnerT<double>(p, o, stream));
    throw BpmError("precision must be 32 or 64");
}

}  // namespace nmk
