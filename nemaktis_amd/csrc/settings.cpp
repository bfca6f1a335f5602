#include "settings.h"

#include <cmath>
#include <algorithm>
#include <exception>
#include <mutex>
#include <regex>
#include <thread>

#include "expr.h"
#include "json.h"

namespace nmk {

namespace {

const double kPi = 3.1415926535897932;

// SettingsLib/BaseSettings.cpp:8-22: a missing section is an error
const Json &section(const Json &parent, const std::string &name) {
    const Json *v = parent.find(name);
    if (!v) throw BpmError("The section \"" + name + "\" is missing from the setting file");
    return *v;
}

// SettingsLib/BaseSettings.h:22-54: typed reads with the section in the message
template <class F>
auto typed(const Json &node, const std::string &sec, const std::string &var, F conv) -> decltype(conv(node)) {
    try {
        return conv(node.at(var));
    } catch (const JsonError &e) {
        throw BpmError("Error when parsing variable \"" + var + "\" in section \"" + sec + "\":\n" + e.what());
    }
}

std::string get_str(const Json &n, const std::string &sec, const std::string &var) {
    return typed(n, sec, var, [](const Json &v) { return v.as_string(); });
}
double get_double(const Json &n, const std::string &sec, const std::string &var) {
    return typed(n, sec, var, [](const Json &v) { return v.as_double(); });
}
long long get_int(const Json &n, const std::string &sec, const std::string &var) {
    return typed(n, sec, var, [](const Json &v) { return v.as_int(); });
}
bool get_bool(const Json &n, const std::string &sec, const std::string &var) {
    return typed(n, sec, var, [](const Json &v) { return v.as_bool(); });
}
std::vector<double> get_vec(const Json &n, const std::string &sec, const std::string &var) {
    return typed(n, sec, var, [](const Json &v) { return v.as_double_vector(); });
}

double eval_index(const std::string &expr, double lambda, const char *name) {
    try {
        double v = eval_index_expression(expr, lambda);
        return v;
    } catch (const ExprError &) {
        throw BpmError(std::string("Wrong expression for ") + name);
    }
}

}  // namespace

Settings parse_settings(const std::string &json_text) {
    Json root;
    try {
        root = Json::parse(json_text);
    } catch (const JsonError &e) {
        throw BpmError(e.what());
    }
    Settings s;
    // Algorithm settings (AlgorithmSettings.cpp:3-22)
    const Json &alg = section(root, "Algorithm settings");
    const Json &gen = section(alg, "General");
    s.lc_field_type = get_str(gen, "General", "LC field type");
    s.results_folder = get_str(gen, "General", "Results folder name");
    if (s.lc_field_type != "Director" && s.lc_field_type != "Q-tensor")
        throw BpmError("Parsing error: \"LC field type\" must be \"Director\" or \"Q-tensor\"");
    const Json &bp = section(alg, "Beam propagation");
    s.n_woodbury = (int)get_int(bp, "Beam propagation", "N Woodbury steps");
    s.nz_substeps = (int)get_int(bp, "Beam propagation", "Number of substeps per slab");

    // Physics settings (PhysicsSettings.cpp:8-88)
    const Json &phys = section(root, "Physics settings");
    const Json &ic = section(phys, "Initial conditions");
    s.beam_profile = get_str(ic, "Initial conditions", "Beam profile");
    s.lc_field_file = get_str(ic, "Initial conditions", "LC field file");
    s.basis = get_str(ic, "Initial conditions", "Basis convention");
    if (s.basis != "XYZ" && s.basis != "YZX" && s.basis != "ZXY")
        throw BpmError("Parsing error: \"Basis convention\" must be \"XYZ\", \"YZX\" or \"ZXY\"");
    const Json &co = section(phys, "Coefficients");
    const std::string cs = "Coefficients";
    s.no_expr = get_str(co, cs, "no");
    s.ne_expr = get_str(co, cs, "ne");
    s.ne_imag_expr = co.find("ne_imag") ? get_str(co, cs, "ne_imag") : "0";
    s.nhost_expr = co.find("nhost") ? get_str(co, cs, "nhost") : "1";
    s.nin_expr = co.find("nin") ? get_str(co, cs, "nin") : "1";
    s.nout_expr = co.find("nout") ? get_str(co, cs, "nout") : "1";
    if (co.find("Wavelengths")) {
        s.wavelengths = get_vec(co, cs, "Wavelengths");
    } else {
        double mean = get_double(co, cs, "Mean wavelength");
        double fwhm = get_double(co, cs, "Spectral FWHM");
        long long nw = get_int(co, cs, "N wavelengths");
        if (nw < 0) throw BpmError("Error when parsing variable \"N wavelengths\" in section \"Coefficients\"");
        if (nw == 1) {
            s.wavelengths.push_back(mean);
        } else {
            for (long long i = 0; i < nw; i++) s.wavelengths.push_back(mean + fwhm * (double(i) / (nw - 1.) - 0.5));
        }
    }
    if (co.find("Wavevectors")) {
        std::vector<double> raw = get_vec(co, cs, "Wavevectors");
        if (raw.size() % 2 != 0)
            throw BpmError("Wrong number of values for the \"Wavevectors\" entry,\nshould be a multiple of 2");
        for (size_t i = 0; i < raw.size() / 2; i++) s.q.emplace_back(raw[2 * i], raw[2 * i + 1]);
    } else {
        double na = get_double(co, cs, "Condenser numerical aperture");
        long long nc = get_int(co, cs, "N radial illumination directions");
        s.q.emplace_back(0.0, 0.0);
        for (long long ir = 1; ir < nc; ir++)
            for (long long iphi = 0; iphi < 6 * ir; iphi++) {
                double phi = iphi * kPi / 3;
                s.q.emplace_back(ir * na / (nc - 1) * std::cos(phi), ir * na / (nc - 1) * std::sin(phi));
            }
    }

    // Postprocessor settings (PostprocessorSettings.cpp:3-19)
    const Json &pp = section(root, "Postprocessor settings");
    const Json &bulk = section(pp, "Bulk output");
    s.bulk_activate = get_bool(bulk, "Bulk output", "Activate");
    s.bulk_name = get_str(bulk, "Bulk output", "Base name");
    const Json &scr = section(pp, "Screen output");
    const std::string ss = "Screen output";
    s.screen_activate = get_bool(scr, ss, "Activate");
    s.screen_name = get_str(scr, ss, "Base name");
    s.iso_thick = get_vec(scr, ss, "Isotropic layer thicknesses");
    s.iso_index = get_vec(scr, ss, "Isotropic layer refractive indices");
    s.z_shift = get_double(scr, ss, "Focalisation z-shift");
    s.NA = get_double(scr, ss, "Numerical aperture");

    // mesh (PythonBackend/bpm_backend.cpp:20-26)
    std::vector<double> sp;
    std::vector<long long> dims;
    try {
        sp = ic.at("Mesh spacings").as_double_vector();
        dims = ic.at("Mesh dimensions").as_int_vector();
    } catch (const JsonError &e) {
        throw BpmError(e.what());
    }
    if (sp.size() < 3 || dims.size() < 3) throw BpmError("\"Mesh spacings\" and \"Mesh dimensions\" need 3 entries");
    for (int i = 0; i < 3; i++)
        if (dims[i] < 1 || dims[i] > (1LL << 30)) throw BpmError("Invalid mesh dimensions");
    s.Nx = (int)dims[0];
    s.Ny = (int)dims[1];
    s.Nz = (int)dims[2];
    s.dx = sp[0];
    s.dy = sp[1];
    s.dz = sp[2];
    return s;
}

// UtilitiesLib/VectorField.cpp:64-113: interface interpolation weights along z
// from a mask formula, by bisection of the formula's zero crossing in each
// slab.  The columns are independent; they are split over host threads as the
// reference splits them over OpenMP threads (the expression tree is read-only).
void interface_weights(const Problem &p, const std::string &formula, std::vector<double> &w, std::string *warning) {
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const int64_t N = (int64_t)Nx * Ny;
    w.assign(2 * N * (int64_t)(Nz - 1), 0.5);
    const Expr e(formula, {"x", "y", "z"});
    ExprProgram prog;
    const bool compiled = e.compile(prog);
    auto columns = [&](int64_t beg, int64_t end) {
        bool failed = false;
        std::vector<double> fplane((size_t)Nz);  // the formula on the planes of one column
        for (int64_t iperp = beg; iperp < end; iperp++) {
            int ix = int(iperp % Nx), iy = int(iperp / Nx);
            double v[3] = {(ix - 0.5 * (Nx - 1)) * p.dx, (iy - 0.5 * (Ny - 1)) * p.dy, 0};
            auto f = [&](double z) {
                v[2] = z;
                return compiled ? expr_run(prog, v) : e.eval(v);
            };
            // each plane bounds two slabs: evaluated once (the same values the
            // slab-by-slab evaluation gives)
            for (int iz = 0; iz < Nz; iz++) fplane[iz] = f((iz - 0.5 * (Nz - 1)) * p.dz);
            for (int iz = 0; iz < Nz - 1; iz++) {
                int64_t i = iperp + N * iz;
                double zi = (iz - 0.5 * (Nz - 1)) * p.dz;
                double zm = zi, zp = (iz + 1 - 0.5 * (Nz - 1)) * p.dz, z = zm;
                double fm = fplane[iz], fp = fplane[iz + 1];
                if (fm * fp <= 0) {
                    bool m0 = p.mask.empty() ? true : p.mask[i] != 0;
                    bool m1 = p.mask.empty() ? true : p.mask[i + N] != 0;
                    if (m0 == m1) failed = true;
                    while (std::fabs(zp - zm) >= 1e-3 * p.dz) {
                        z = zm;
                        fm = f(z);
                        z = 0.5 * (zp + zm);
                        double fz = f(z);
                        if (fz == 0) break;
                        if (fz * fm < 0) zp = z;
                        else zm = z;
                    }
                    w[2 * i] = (z - zi) / p.dz;
                    w[2 * i + 1] = 1 - w[2 * i];
                }
            }
        }
        return failed;
    };
    const int64_t nthreads = std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), N / 256));
    bool failed = false;
    if (nthreads == 1) {
        failed = columns(0, N);
    } else {
        std::vector<std::thread> pool;
        std::vector<char> flags((size_t)nthreads, 0);
        std::exception_ptr first_error;
        std::mutex mu;
        for (int64_t t = 0; t < nthreads; t++)
            pool.emplace_back([&, t] {
                try {
                    flags[(size_t)t] = columns(N * t / nthreads, N * (t + 1) / nthreads) ? 1 : 0;
                } catch (...) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (!first_error) first_error = std::current_exception();
                }
            });
        for (auto &th : pool) th.join();
        if (first_error) std::rethrow_exception(first_error);
        for (char c : flags) failed = failed || c;
    }
    if (failed) {
        // the reference prints this and falls back to the default weights
        if (warning) *warning = kMaskMismatchNotice;
        w.clear();
    }
}

const char *const kMaskMismatchNotice =
    "Mask values do not corresponds to the given formula, will ignore interpolation correction";

void prepare_mask(Problem &p, std::string *warning) {
    if (!p.mask_src) return;
    const int64_t nv = p.voxels();
    p.mask.resize((size_t)nv);
    const double *src = p.mask_src;
    uint8_t *dst = p.mask.data();
    auto flags = [src, dst](int64_t beg, int64_t end) {
        for (int64_t i = beg; i < end; i++) dst[i] = src[i] >= 0 ? 1 : 0;
    };
    const int64_t nthreads = std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), nv >> 22));
    if (nthreads == 1) {
        flags(0, nv);
    } else {
        std::vector<std::thread> pool;
        for (int64_t t = 0; t < nthreads; t++) pool.emplace_back(flags, nv * t / nthreads, nv * (t + 1) / nthreads);
        for (auto &th : pool) th.join();
    }
    if (!p.mask_formula.empty()) {
        try {
            const Expr e(p.mask_formula, {"x", "y", "z"});
            p.weights_on_device = e.compile(p.weight_prog);
            if (!p.weights_on_device) interface_weights(p, p.mask_formula, p.weights, warning);
        } catch (const ExprError &e) {
            throw BpmError(std::string("Wrong mask formula: ") + e.what());
        }
    }
    p.mask_src = nullptr;
}

Problem build_problem(const Settings &s, const double *lc, int64_t n_lc, const double *mask, int64_t n_mask,
                      const char *mask_formula, int64_t n_E, std::string *warning, bool defer_mask) {
    Problem p;
    p.Nx = s.Nx;
    p.Ny = s.Ny;
    p.Nz = s.Nz;
    p.dx = s.dx;
    p.dy = s.dy;
    p.dz = s.dz;
    const int64_t nv = p.voxels();
    // UtilitiesLib/VectorField.cpp:26-62: field dimension from the array size
    int64_t fd = nv > 0 ? n_lc / nv : 0;
    if (n_lc != nv * fd) {
        if (mask) throw BpmError("Wrong dimension for the user pointer array");
        throw BpmError("Wrong dimension for the user pointer array (VectorField)");
    }
    if (mask) {
        if (n_mask != nv) throw BpmError("Wrong dimension for the mask pointer array");
        p.mask_src = mask;
        p.mask_formula = mask_formula ? mask_formula : "";
    }
    if (lc == nullptr && n_lc > 0) throw BpmError("Null LC field pointer");
    p.lc = lc;
    if (mask && !p.mask_formula.empty()) {
        try {
            Expr(p.mask_formula, {"x", "y", "z"});  // syntax check now, evaluation with the mask
        } catch (const ExprError &e) {
            throw BpmError(std::string("Wrong mask formula: ") + e.what());
        }
    }
    if (mask && !defer_mask) prepare_mask(p, warning);
    // PythonBackend/bpm_backend.cpp:29-31
    if (fd != 3 && fd != 6) throw BpmError("Unexpected dimension for the LC orientational field, should be 3 or 6.");
    p.field_dim = (int)fd;
    p.wl = s.wavelengths;
    for (const auto &q : s.q) {
        p.qx.push_back(q.first);
        p.qy.push_back(q.second);
    }
    // BeamPropagationLib/OpticalFieldCollection.cpp:38-39
    if (n_E != p.n_out_values())
        throw BpmError("Wrong dimension for the user pointer array (OpticalFieldCollection)");
    // BeamPropagationLib/BPMIteration.cpp:30-46
    if (s.beam_profile == "UniformBeam") {
        p.beam = BeamKind::Uniform;
    } else if (s.beam_profile == "None") {
        p.beam = BeamKind::None;
    } else {
        std::smatch m;
        static const std::regex re("GaussianBeam\\(([0-9]+\\.?[0-9]*)\\)");
        if (std::regex_match(s.beam_profile, m, re)) {
            p.beam = BeamKind::Gaussian;
            p.waist = std::stod(m[1]);
        } else {
            throw BpmError("Unrecognised beam profile");
        }
    }
    p.n_woodbury = s.n_woodbury;
    if (p.n_woodbury < 0) throw BpmError("\"N Woodbury steps\" must be non-negative");
    // PostprocessorLib/ScreenOutput.cpp:28-38
    if (s.iso_thick.size() != s.iso_index.size()) throw BpmError("Inconsistent number of isotropic layers");
    p.iso_thick = s.iso_thick;
    p.iso_index = s.iso_index;
    p.nout_ref = eval_index(s.nout_expr, 0.6, "nout");
    p.z_foc = s.z_shift;
    for (size_t i = 0; i < p.iso_thick.size(); i++) p.z_foc -= p.iso_thick[i] * p.nout_ref / p.iso_index[i];
    p.NA = s.NA;
    // PythonBackend/bpm_backend.cpp:44-47: the bulk fields go to
    // "<Results folder name>/<Base name>.vti" (the folder is used unless "" or ".")
    p.bulk = s.bulk_activate;
    if (p.bulk) {
        const bool dir = !s.results_folder.empty() && s.results_folder != ".";
        p.bulk_dir = dir ? s.results_folder : "";
        p.bulk_path = (dir ? s.results_folder + "/" : std::string()) + s.bulk_name + ".vti";
    }
    // PhysicsLib/PhysicsCoefficients.cpp:27-103, evaluated for every wavelength up front
    for (double l : p.wl) {
        p.no.push_back(eval_index(s.no_expr, l, "no"));
        p.ne.push_back(eval_index(s.ne_expr, l, "ne"));
        p.ne_imag.push_back(eval_index(s.ne_imag_expr, l, "ne_imag"));
        p.nhost.push_back(eval_index(s.nhost_expr, l, "nhost"));
        p.nin.push_back(eval_index(s.nin_expr, l, "nin"));
        p.nout.push_back(eval_index(s.nout_expr, l, "nout"));
        p.nout_fresnel.push_back(p.iso_index.empty() ? p.nout.back() : p.iso_index[0]);
    }
    return p;
}

}  // namespace nmk
