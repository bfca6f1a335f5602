// Host code of the library under AddressSanitizer and UBSan (built by
// tests/test_host_sanitizers.py with g++, no GPU): settings JSON (valid,
// truncated, malformed, deeply nested), expressions (valid, malformed,
// nested past the parser bound) with the expression tree and the compiled
// program compared, host interface weights, and a bulk VTI file.
#include <cstdio>
#include <string>
#include <vector>
#include "settings.h"
#include "expr.h"
#include "vti.h"
using namespace nmk;
int main(int argc, char **argv) {
    const std::string bulk_path = argc > 1 ? argv[1] : "host_driver_bulk.vti";
    int mismatches = 0;
    const char *good = R"({"Algorithm settings":{"General":{"LC field type":"Director","Results folder name":""},"Beam propagation":{"N Woodbury steps":2,"Number of substeps per slab":1}},"Physics settings":{"Initial conditions":{"Beam profile":"UniformBeam","LC field file":"","Mesh dimensions":[8,6,4],"Mesh spacings":[0.1,0.1,0.1],"Basis convention":"XYZ"},"Coefficients":{"no":"1.5","ne":"1.7+0.01/lambda^2","nhost":"1.5","nin":"1","nout":"1","Wavelengths":[0.5,0.6],"Wavevectors":[0,0,0.1,0.05]}},"Postprocessor settings":{"Bulk output":{"Activate":false,"Base name":""},"Screen output":{"Activate":true,"Base name":"","Isotropic layer thicknesses":[0.1],"Isotropic layer refractive indices":[1.4],"Focalisation z-shift":0,"Numerical aperture":0.4}}})";
    std::vector<std::string> inputs = {good, "", "{", "[1,2", "{\"a\":\"\\u12\"}", "{\"a\":1e999999}", "{\"a\":[[[[[[[[[[[[[[[[[[[[]]]]]]]]]]]]]]]]]]]]}",
                                       std::string(good).substr(0, 200), "{\"Physics settings\":{}}", "nul", "{\"a\":\"\\\"}"};
    int ok = 0, bad = 0;
    for (auto &s : inputs) {
        try { Settings st = parse_settings(s); ok++; (void)st; } catch (const std::exception &e) { bad++; }
    }
    std::vector<std::string> exprs = {"1+2*3", "sin(x)^2+cos(y)^2", "(x>0 ? 1 : -1)*z", "min(x,y,z)+max(1,2)+avg(1,2,3)+sum(4)",
        "-2^2", "2^-1", "1/0", "((((((((((((((((((((x))))))))))))))))))))", "x+", "sin(", "foo(1)", "1 ? 2", "_pi*_e", "atan2(1,2)",
        std::string(5000, '(') + "1" + std::string(5000, ')'), "1e308*10", "abs(-3)+sign(-2)+rint(2.5)"};
    for (auto &e : exprs) {
        try {
            Expr ex(e, {"x", "y", "z"});
            double v[3] = {0.3, -0.2, 0.1};
            double a = ex.eval(v);
            ExprProgram p;
            if (ex.compile(p)) { double b = expr_run(p, v); if (!(a == b || (a != a && b != b))) { std::printf("MISMATCH %s %g %g\n", e.c_str(), a, b); mismatches++; } }
            ok++;
        } catch (const std::exception &) { bad++; }
    }
    Settings st = parse_settings(good);
    std::vector<double> lc(8 * 6 * 4 * 3, 0.0), mask(8 * 6 * 4, 1.0);
    for (size_t i = 0; i < mask.size(); i++) mask[i] = (i % 7) < 3 ? -1.0 : 1.0;
    const int64_t nE = 2 * 2 * 4 * 8 * 6;
    for (const char *f : {"0.05-z", "x^2+y^2+z^2-0.04", "z-0.3*sin(2*x)", ""}) {
        std::string w;
        try {
            Problem p = build_problem(st, lc.data(), (int64_t)lc.size(), mask.data(), (int64_t)mask.size(), f, nE, &w);
            std::vector<double> wts;
            if (*f) interface_weights(p, f, wts, &w);
            ok++;
        } catch (const std::exception &e) { bad++; std::printf("problem %s: %s\n", f, e.what()); }
    }
    {
        BulkGeometry bg;
        bg.Nx = 7; bg.Ny = 5; bg.Nz = 3; bg.dx = bg.dy = bg.dz = 0.1; bg.wl = {0.5, 0.6}; bg.qx = {0, 0.1}; bg.qy = {0, 0};
        const size_t blk = (size_t)bg.Nz * 2 * bg.Nx * bg.Ny;
        write_bulk_vti(bulk_path, "", bg, [&](size_t k, c64 *dst) {
            for (size_t i = 0; i < blk; i++) dst[i] = mk(double(k), double(i));
        });
        ok++;
    }
    std::printf("ok %d refused %d mismatches %d\n", ok, bad, mismatches);
    return mismatches ? 1 : 0;
}
