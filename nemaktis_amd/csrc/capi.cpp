// extern "C" entry points of libnemaktis_bpm.so (declared in include/nemaktis_bpm.h).
#include "../../include/nemaktis_bpm.h"
#include "../../include/nemaktis_view.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>

#include "engine.h"
#include "expr.h"
#include "noalloc.h"
#include "server.h"
#include "settings.h"
#include "viewer.h"

using namespace nmk;

namespace {

void set_err(char *buf, size_t len, const std::string &msg) {
    if (!buf || len == 0) return;
    size_t n = std::min(len - 1, msg.size());
    std::memcpy(buf, msg.data(), n);
    buf[n] = '\0';
}

Options to_options(const nmk_bpm_options *o) {
    Options r;
    if (o) {
        r.precision = o->precision;
        r.wl_batch = o->wl_batch;
        r.use_graph = o->use_graph;
        r.device = o->device;
        r.profile = o->profile;
        if (o->x_chunks > 0 && o->x_rows > 0 && o->y_chunks > 0 && o->y_cols > 0) {
            r.tuning.x_chunks = o->x_chunks;
            r.tuning.x_rows = o->x_rows;
            r.tuning.y_chunks = o->y_chunks;
            r.tuning.y_cols = o->y_cols;
            if (o->item_order > 0) r.tuning.order = (o->item_order - 1) & 15;
            r.tuning.lean = o->lean ? 1 : 0;
            r.tuning.y_pair = o->y_pair == 1 || o->y_pair == 2 ? o->y_pair : 0;
            r.tuning.x_both = o->x_both & 3;
            r.tuning.coef_serial = o->coef_serial ? 1 : 0;
        }
    }
    return r;
}

void fill_stats(nmk_bpm_stats *out, const Stats &s) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->setup_ms = s.setup_ms;
    out->march_ms = s.march_ms;
    out->output_ms = s.output_ms;
    out->total_ms = s.total_ms;
    out->updates = s.updates;
    out->updates_per_s = s.total_ms > 0 ? s.updates / (s.total_ms * 1e-3) : 0;
    out->footprint_bytes = s.footprint_bytes;
    for (int i = 0; i < KF_COUNT && i < 10; i++) {
        out->family_ms[i] = s.family_ms[i];
        out->family_launches[i] = s.family_launches[i];
    }
}

// maps exceptions to status codes and messages; on success the message
// buffer receives the notice (if any) the reference would have printed
template <class F>
int guarded(char *err, size_t len, F &&f, const std::string *notice = nullptr) {
    try {
        f();
        set_err(err, len, notice ? *notice : std::string());
        return NMK_BPM_OK;
    } catch (const BpmError &e) {
        set_err(err, len, e.what());
        return NMK_BPM_EINPUT;
    } catch (const std::bad_alloc &) {
        set_err(err, len, "host out of memory");
        return NMK_BPM_ENOMEM;
    } catch (const ScratchExhausted &e) {
        set_err(err, len, e.what());
        return NMK_BPM_ENOMEM;
    } catch (const std::runtime_error &e) {
        std::string m = e.what();
        set_err(err, len, m);
        if (m.find("out of memory") != std::string::npos || m.find("hipErrorOutOfMemory") != std::string::npos)
            return NMK_BPM_ENOMEM;
        return m.rfind("HIP", 0) == 0 ? NMK_BPM_EDEVICE : NMK_BPM_EINTERNAL;
    } catch (const std::exception &e) {
        set_err(err, len, e.what());
        return NMK_BPM_EINTERNAL;
    } catch (...) {
        set_err(err, len, "Unknown exception!");
        return NMK_BPM_EINTERNAL;
    }
}

Problem problem_from(const char *json, const char *formula, const double *lc, int64_t n_lc, const double *mask,
                     int64_t n_mask, int64_t n_E, std::string *warning) {
    if (!json) throw BpmError("null settings string");
    Settings s = parse_settings(json);
    return build_problem(s, lc, n_lc, mask, n_mask, mask ? (formula ? formula : "") : nullptr, n_E, warning);
}

ViewGeometry view_geometry(const nmk_view_geometry *g) {
    if (!g) throw BpmError("null geometry");
    ViewGeometry v;
    v.Nx = g->Nx;
    v.Ny = g->Ny;
    v.Lx = g->Lx;
    v.Ly = g->Ly;
    if (g->n_wavelengths < 1 || !g->wavelengths) throw BpmError("at least one wavelength is required");
    v.wl.assign(g->wavelengths, g->wavelengths + g->n_wavelengths);
    if (g->n_out) v.n_out.assign(g->n_out, g->n_out + g->n_wavelengths);
    if (g->n_wavevectors > 0) {
        if (!g->wavevectors) throw BpmError("null wavevectors");
        for (int q = 0; q < g->n_wavevectors; q++) {
            v.qx.push_back(g->wavevectors[2 * q]);
            v.qy.push_back(g->wavevectors[2 * q + 1]);
        }
    }
    v.max_NA_objective = g->max_NA_objective;
    v.max_NA_condenser = g->max_NA_condenser;
    v.Nr = g->N_radial_wavevectors;
    v.koehler_1D = g->koehler_1D != 0;
    v.zfoc_NA_corr = g->zfoc_NA_corr;
    v.device = g->device;
    return v;
}

ViewOptics view_optics(const nmk_view_optics *o) {
    ViewOptics r;
    if (!o) return r;
    r.polariser = o->polariser != 0;
    r.analyser = o->analyser != 0;
    r.polariser_angle = o->polariser_angle;
    r.analyser_angle = o->analyser_angle;
    r.lower_waveplate = o->lower_waveplate;
    r.upper_waveplate = o->upper_waveplate;
    for (int k : {r.lower_waveplate, r.upper_waveplate})
        if (k < NMK_VIEW_NO_WAVEPLATE || k > NMK_VIEW_TINT_SENSITIVE)
            throw BpmError("unknown waveplate type " + std::to_string(k));
    r.lower_waveplate_angle = o->lower_waveplate_angle;
    r.upper_waveplate_angle = o->upper_waveplate_angle;
    r.intensity = o->intensity;
    r.NA_condenser = o->NA_condenser;
    r.grayscale = o->grayscale != 0;
    r.gamma = o->gamma != 0;
    return r;
}

}  // namespace

struct nmk_view {
    std::unique_ptr<Viewer> v;
};

struct nmk_bpm_ctx {
    Problem problem;
    std::vector<double> lc_copy;
    std::unique_ptr<Runner> runner;
};

extern "C" {

void nmk_bpm_default_options(nmk_bpm_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = 64;
    o->wl_batch = 1;
    o->use_graph = 0;
    o->device = -1;
}

int nmk_bpm_run(const char *json, const char *formula, const double *lc, int64_t n_lc, const double *mask,
                int64_t n_mask, double *E, int64_t n_E, const nmk_bpm_options *opts, nmk_bpm_stats *stats, char *err,
                size_t errlen) {
    std::string warning;
    return guarded(err, errlen, [&] {
        Problem p = problem_from(json, formula, lc, n_lc, mask, n_mask, n_E, &warning);
        Options o = to_options(opts);
        // a one-shot run marches once: a captured graph would be replayed
        // once and only adds the capture, so one-shot runs launch eagerly
        // (use_graph is honoured by contexts, which replay their graphs)
        o.use_graph = 0;
        auto r = make_runner(p, o);
        if (!r->notice().empty()) warning = r->notice();
        if (p.beam == BeamKind::None) r->upload_fields(reinterpret_cast<const c64 *>(E));
        Stats st;
        r->propagate(&st);
        r->download(reinterpret_cast<c64 *>(E));
        fill_stats(stats, st);
    }, &warning);
}

int nmk_bpm_run_backend_without_mask(const char *json, const double *lc, int64_t n_lc, double *E, int64_t n_E,
                                     char *err, size_t errlen) {
    nmk_bpm_options o;
    nmk_bpm_default_options(&o);
    return nmk_bpm_run(json, nullptr, lc, n_lc, nullptr, 0, E, n_E, &o, nullptr, err, errlen);
}

int nmk_bpm_run_backend_with_mask(const char *json, const char *formula, const double *lc, int64_t n_lc,
                                  const double *mask, int64_t n_mask, double *E, int64_t n_E, char *err,
                                  size_t errlen) {
    if (!mask) {
        set_err(err, errlen, "Wrong dimension for the mask pointer array");
        return NMK_BPM_EINPUT;
    }
    nmk_bpm_options o;
    nmk_bpm_default_options(&o);
    return nmk_bpm_run(json, formula ? formula : "", lc, n_lc, mask, n_mask, E, n_E, &o, nullptr, err, errlen);
}

int nmk_bpm_run_noalloc(const char *json, const double *lc_dev, int64_t n_lc, const double *mask_dev,
                        int64_t n_mask, const double *weights_dev, int64_t n_weights, double *E_dev, int64_t n_E,
                        int32_t precision, void *stream, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!json) throw BpmError("null settings string");
        if (!lc_dev || !E_dev) throw BpmError("null device array");
        if (precision != 32 && precision != 64) throw BpmError("precision must be 32 or 64");
        const Settings s = parse_settings(json);
        // the checks of nmk_bpm_run; the mask stays on the device (no formula:
        // its interface weights come precomputed in weights_dev)
        Problem p = build_problem(s, lc_dev, n_lc, mask_dev, mask_dev ? n_mask : 0, mask_dev ? "" : nullptr, n_E,
                                  nullptr, /*defer_mask=*/true);
        p.mask_src = nullptr;
        p.lc = nullptr;
        if (weights_dev && n_weights != 2 * (int64_t)p.Nx * p.Ny * (p.Nz - 1))
            throw BpmError("the weights array needs 2*Nx*Ny*(Nz-1) values");
        if (p.bulk) throw BpmError("the no-allocation entry point does not write bulk output");
        DeviceIO io;
        io.lc = lc_dev;
        io.mask = mask_dev;
        io.weights = weights_dev;
        io.E = reinterpret_cast<c64 *>(E_dev);
        run_noalloc(p, precision, io, static_cast<hipStream_t>(stream));
    });
}

nmk_bpm_ctx *nmk_bpm_create(const char *json, const char *formula, const double *lc, int64_t n_lc,
                            const double *mask, int64_t n_mask, int64_t n_E, const nmk_bpm_options *opts, char *err,
                            size_t errlen) {
    nmk_bpm_ctx *ctx = nullptr;
    std::string warning;
    int rc = guarded(err, errlen, [&] {
        std::unique_ptr<nmk_bpm_ctx> c(new nmk_bpm_ctx);
        c->problem = problem_from(json, formula, lc, n_lc, mask, n_mask, n_E, &warning);
        c->runner = make_runner(c->problem, to_options(opts));
        if (!c->runner->notice().empty()) warning = c->runner->notice();
        c->problem.lc = nullptr;  // the director now lives in HBM
        ctx = c.release();
    }, &warning);
    return rc == NMK_BPM_OK ? ctx : nullptr;
}

int nmk_bpm_upload_fields(nmk_bpm_ctx *ctx, const double *E, int64_t n_E, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!ctx) throw BpmError("null context");
        if (n_E != ctx->problem.n_out_values()) throw BpmError("Wrong dimension for the raw fields array");
        ctx->runner->upload_fields(reinterpret_cast<const c64 *>(E));
    });
}

int nmk_bpm_propagate(nmk_bpm_ctx *ctx, nmk_bpm_stats *stats, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!ctx) throw BpmError("null context");
        Stats st;
        ctx->runner->propagate(&st);
        fill_stats(stats, st);
    });
}

int nmk_bpm_set_tuning(nmk_bpm_ctx *ctx, const nmk_bpm_options *opts, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!ctx) throw BpmError("null context");
        ctx->runner->retune(to_options(opts).tuning);
    });
}

int nmk_bpm_download(nmk_bpm_ctx *ctx, double *E, int64_t n_E, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!ctx) throw BpmError("null context");
        if (n_E != ctx->problem.n_out_values()) throw BpmError("Wrong dimension for the raw fields array");
        ctx->runner->download(reinterpret_cast<c64 *>(E));
    });
}

void nmk_bpm_destroy(nmk_bpm_ctx *ctx) { delete ctx; }

int nmk_bpm_estimate_footprint(const char *json, int64_t n_lc, int has_mask, const nmk_bpm_options *opts,
                               int64_t *bytes, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        Settings s = parse_settings(json ? json : "");
        Problem p;
        p.Nx = s.Nx;
        p.Ny = s.Ny;
        p.Nz = s.Nz;
        int64_t nv = p.voxels();
        p.field_dim = nv > 0 ? (int)(n_lc / nv) : 3;
        p.wl = s.wavelengths;
        for (auto &q : s.q) {
            p.qx.push_back(q.first);
            p.qy.push_back(q.second);
        }
        p.bulk = s.bulk_activate;
        int64_t b = estimate_footprint(p, to_options(opts));
        if (has_mask) {
            // the mask and (upper bound) the interface weights of a mask formula
            auto al = [](int64_t v) { return (v + (int64_t)kAllocAlign - 1) / (int64_t)kAllocAlign * (int64_t)kAllocAlign; };
            b += al(nv) + al(2 * (int64_t)p.Nx * p.Ny * std::max(0, p.Nz - 1) * 8) + al(sizeof(int));
        }
        if (bytes) *bytes = b;
    });
}

namespace {

int interface_weights_call(const char *json, const char *formula, const double *mask, int64_t n_mask,
                           double *weights, int64_t n_weights, char *err, size_t errlen, bool device) {
    std::string warning;
    return guarded(err, errlen, [&] {
        if (!json) throw BpmError("null settings string");
        if (!formula || !mask) throw BpmError("a mask and a mask formula are required");
        const Settings s = parse_settings(json);
        Problem p;
        p.Nx = s.Nx;
        p.Ny = s.Ny;
        p.Nz = s.Nz;
        p.dx = s.dx;
        p.dy = s.dy;
        p.dz = s.dz;
        const int64_t nv = p.voxels();
        if (n_mask != nv) throw BpmError("Wrong dimension for the mask pointer array");
        const int64_t nw = 2 * (int64_t)p.Nx * p.Ny * (p.Nz - 1);
        if (!weights || n_weights != nw) throw BpmError("the weights array needs 2*Nx*Ny*(Nz-1) values");
        p.mask.resize((size_t)nv);
        for (int64_t i = 0; i < nv; i++) p.mask[i] = mask[i] >= 0 ? 1 : 0;
        std::vector<double> w;
        try {
            if (device) {
                const Expr e(formula, {"x", "y", "z"});
                p.weights_on_device = e.compile(p.weight_prog);
                if (!p.weights_on_device) throw BpmError("the mask formula is too long for a device program");
                device_interface_weights(p, w, &warning);
            } else {
                interface_weights(p, formula, w, &warning);
            }
        } catch (const ExprError &e) {
            throw BpmError(std::string("Wrong mask formula: ") + e.what());
        }
        if (w.empty()) std::fill(weights, weights + nw, 0.5);
        else std::copy(w.begin(), w.end(), weights);
    }, &warning);
}

}  // namespace

int nmk_bpm_interface_weights(const char *json, const char *formula, const double *mask, int64_t n_mask,
                              double *weights, int64_t n_weights, char *err, size_t errlen) {
    return interface_weights_call(json, formula, mask, n_mask, weights, n_weights, err, errlen, false);
}

int nmk_bpm_interface_weights_device(const char *json, const char *formula, const double *mask, int64_t n_mask,
                                     double *weights, int64_t n_weights, char *err, size_t errlen) {
    return interface_weights_call(json, formula, mask, n_mask, weights, n_weights, err, errlen, true);
}

int nmk_bpm_eval_expression(const char *expr, double lambda, double *value, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!expr) throw BpmError("null expression");
        try {
            double v = eval_index_expression(expr, lambda);
            if (value) *value = v;
        } catch (const ExprError &e) {
            throw BpmError(e.what());
        }
    });
}

const char *nmk_bpm_version(void) { return "nemaktis-amd bpm 0.2 (gfx950)"; }

/* ---- serving loop ---------------------------------------------------------- */

void nmk_srv_default_config(nmk_srv_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->device = -1;
    c->max_active = 4;
    c->hbm_budget = 0;
    c->max_bypass = 8;
}

nmk_srv *nmk_srv_create(const nmk_srv_config *cfg, char *err, size_t errlen) {
    nmk_srv *srv = nullptr;
    guarded(err, errlen, [&] {
        nmk_srv_config d;
        nmk_srv_default_config(&d);
        if (!cfg) cfg = &d;
        ServerConfig c;
        c.device = cfg->device;
        c.max_active = cfg->max_active;
        c.hbm_budget = cfg->hbm_budget;
        c.max_bypass = cfg->max_bypass;
        srv = reinterpret_cast<nmk_srv *>(new Server(c));
    });
    return srv;
}

int64_t nmk_srv_submit(nmk_srv *srv, const char *json, const char *formula, const double *lc, int64_t n_lc,
                       const double *mask, int64_t n_mask, double *E, int64_t n_E, const nmk_bpm_options *opts,
                       char *err, size_t errlen) {
    int64_t id = -1;
    const int rc = guarded(err, errlen, [&] {
        if (!srv) throw BpmError("null server");
        if (!json) throw BpmError("null settings string");
        id = reinterpret_cast<Server *>(srv)->submit(json, formula ? formula : "", lc, n_lc, mask, n_mask,
                                                      reinterpret_cast<c64 *>(E), n_E, to_options(opts));
    });
    return rc == NMK_BPM_OK ? id : -(int64_t)rc;
}

int nmk_srv_wait(nmk_srv *srv, int64_t id, double timeout_ms, nmk_srv_job_info *info) {
    if (!srv) return -1;
    JobInfo ji;
    const int r = reinterpret_cast<Server *>(srv)->wait(id, timeout_ms, &ji);
    if (info && r >= 0) {
        std::memset(info, 0, sizeof(*info));
        info->state = ji.state;
        info->status = ji.status;
        info->t_submit_ms = ji.t_submit_ms;
        info->t_admit_ms = ji.t_admit_ms;
        info->t_end_ms = ji.t_end_ms;
        info->device_ms = ji.device_ms;
        info->updates = ji.updates;
        info->footprint_bytes = ji.footprint;
        set_err(info->error, sizeof(info->error), ji.error);
        info->t_ready_ms = ji.t_ready_ms;
        set_err(info->notice, sizeof(info->notice), ji.notice);
    }
    return r;
}

int nmk_srv_cancel(nmk_srv *srv, int64_t id) {
    if (!srv) return -1;
    bool known = false;
    const bool ok = reinterpret_cast<Server *>(srv)->cancel(id, &known);
    return !known ? -1 : ok ? 0 : 1;
}

int nmk_srv_stats(nmk_srv *srv, nmk_srv_counters *out) {
    if (!srv || !out) return -1;
    const ServerStats s = reinterpret_cast<Server *>(srv)->stats();
    std::memset(out, 0, sizeof(*out));
    out->submitted = s.submitted;
    out->completed = s.completed;
    out->failed = s.failed;
    out->cancelled = s.cancelled;
    out->running = s.running;
    out->queued = s.queued;
    out->arena_bytes = s.arena_bytes;
    out->arena_used_bytes = s.arena_used;
    out->arena_peak_bytes = s.arena_peak;
    out->peak_active = s.peak_active;
    out->now_ms = reinterpret_cast<Server *>(srv)->now_ms();
    return 0;
}

void nmk_srv_destroy(nmk_srv *srv) { delete reinterpret_cast<Server *>(srv); }

/* ---- microscope imaging ---------------------------------------------------- */

void nmk_view_default_optics(nmk_view_optics *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->polariser = 1;
    o->analyser = 1;
    o->polariser_angle = 0;
    o->analyser_angle = 90;
    o->intensity = 1;
    o->gamma = 1;
}

nmk_view *nmk_view_create(const nmk_view_geometry *geometry, const double *E, int64_t n_E, char *err,
                          size_t errlen) {
    nmk_view *out = nullptr;
    guarded(err, errlen, [&] {
        std::unique_ptr<nmk_view> h(new nmk_view);
        h->v.reset(new Viewer(view_geometry(geometry)));
        if (E) {
            if (n_E != h->v->geometry().n_values()) throw BpmError("Wrong shape for the optical field ndarray");
            h->v->set_fields(reinterpret_cast<const c64 *>(E));
        }
        out = h.release();
    });
    return out;
}

nmk_view *nmk_view_from_ctx(nmk_bpm_ctx *ctx, const nmk_view_geometry *geometry, char *err, size_t errlen) {
    nmk_view *out = nullptr;
    guarded(err, errlen, [&] {
        if (!ctx) throw BpmError("null context");
        const Problem &p = ctx->problem;
        ViewGeometry g;
        if (geometry) {
            g = view_geometry(geometry);
        } else {
            // as LightPropagator._bpm_propagation builds its OpticalFields
            g.Nx = p.Nx;
            g.Ny = p.Ny;
            g.Lx = p.dx * std::max(1, p.Nx - 1);
            g.Ly = p.dy * std::max(1, p.Ny - 1);
            g.wl = p.wl;
            g.n_out = p.nout;
            g.qx = p.qx;
            g.qy = p.qy;
            g.max_NA_objective = p.NA;
            const double n0 = p.nout_ref;
            for (size_t i = 0; i < p.iso_thick.size(); i++) {
                const double nk = p.iso_index[i], ek = p.iso_thick[i];
                g.zfoc_NA_corr += 3 * n0 * ek / (14 * nk) * (1 / (n0 * n0) - 1 / (nk * nk));
            }
            // rings and condenser aperture from the wavevectors, as the VTI import
            const int Nq = (int)p.qx.size();
            g.Nr = (int)std::lround((1 + std::sqrt((4.0 * Nq - 1) / 3)) / 2);
            g.max_NA_condenser = std::sqrt(p.qx.back() * p.qx.back() + p.qy.back() * p.qy.back());
        }
        if (g.Nx != p.Nx || g.Ny != p.Ny || g.wl.size() != p.wl.size() ||
            (!g.qx.empty() && g.qx.size() != p.qx.size()))
            throw BpmError("the geometry does not match the context's fields");
        std::unique_ptr<nmk_view> h(new nmk_view);
        h->v.reset(new Viewer(g));
        if (h->v->geometry().n_values() != p.n_out_values())
            throw BpmError("the geometry does not match the context's fields");
        h->v->set_fields_device(ctx->runner->device_output(), ctx->runner->stream());
        out = h.release();
    });
    return out;
}

int nmk_view_set_fields(nmk_view *v, const double *E, int64_t n_E, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!v) throw BpmError("null view");
        if (!E || n_E != v->v->geometry().n_values()) throw BpmError("Wrong shape for the optical field ndarray");
        v->v->set_fields(reinterpret_cast<const c64 *>(E));
    });
}

int nmk_view_focus(nmk_view *v, double z_focus, double NA_objective, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!v) throw BpmError("null view");
        v->v->focus(z_focus, NA_objective);
    });
}

int nmk_view_get_focused(nmk_view *v, double *E, int64_t n_E, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!v) throw BpmError("null view");
        if (!E || n_E != v->v->geometry().n_values()) throw BpmError("Wrong shape for the optical field ndarray");
        v->v->download_focused(reinterpret_cast<c64 *>(E));
    });
}

int nmk_view_render(nmk_view *v, const nmk_view_optics *optics, const double *cmf, int64_t n_cmf, char *err,
                    size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!v) throw BpmError("null view");
        if (cmf && n_cmf != 3 * (int64_t)v->v->geometry().wl.size())
            throw BpmError("the colour matching table needs 3 values per wavelength");
        nmk_view_optics d;
        nmk_view_default_optics(&d);
        v->v->render(view_optics(optics ? optics : &d), cmf);
    });
}

int nmk_view_get_spectrum(nmk_view *v, double *spec, int64_t n, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!v) throw BpmError("null view");
        const ViewGeometry &g = v->v->geometry();
        if (!spec || n != (int64_t)g.Nx * g.Ny * (int64_t)g.wl.size())
            throw BpmError("the spectrum array needs Ny * Nx * N_wavelengths values");
        v->v->download_spectrum(spec);
    });
}

int nmk_view_get_image(nmk_view *v, double *img, int64_t n, char *err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!v) throw BpmError("null view");
        const ViewGeometry &g = v->v->geometry();
        if (!img || n != (int64_t)g.Nx * g.Ny * 3) throw BpmError("the image array needs Ny * Nx * 3 values");
        v->v->download_image(img);
    });
}

int nmk_view_state(const nmk_view *v, double *z_focus, double *NA_objective) {
    if (!v) return NMK_BPM_EINPUT;
    if (z_focus) *z_focus = v->v->z_focus();
    if (NA_objective) *NA_objective = v->v->NA_objective();
    return NMK_BPM_OK;
}

int nmk_view_timings(const nmk_view *v, double *ms, int32_t n) {
    if (!v || !ms) return 0;
    const int k = std::min<int>(n, Viewer::T_COUNT);
    for (int i = 0; i < k; i++) ms[i] = v->v->timings()[i];
    return k;
}

void nmk_view_destroy(nmk_view *v) { delete v; }

}  // extern "C"
