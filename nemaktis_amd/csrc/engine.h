// Device engine of the bpm-solver march: owns the HBM buffers of one problem
// and runs BPMIteration::propagate_fields + ScreenOutput::apply_no_export
// (BeamPropagationSolver/src/BeamPropagationLib/BPMIteration.cpp:53-177,
// PostprocessorLib/ScreenOutput.cpp:184-224) on one GPU.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device_common.h"
#include "settings.h"

namespace nmk {

// launch shapes of the line solvers: chunks per line and lines per block
struct Tuning {
    int x_chunks = 0, x_rows = 0;  // x lines: chunks per row, rows per block
    int y_chunks = 0, y_cols = 0;  // y lines: chunks per column, columns per block
    int order = -1;                // XCD work order of the solvers (bits, see launch_xsolve); -1: by batch size
    int lean = 0;                  // 1: register-lean solver variants (see launch_xsolve)
    int y_pair = 0;                // 1: y-line blocks hold both polarisations, 2: both components
    int x_both = 0;                // x-line blocks hold both field components: bit 0 first launch, bit 1 Neumann
    int coef_serial = 0;           // 1: slab coefficients on the march stream (no second stream)
};
Tuning default_tuning(int precision, int Mx, int My, int nprops);

// source of a runner's device memory: hipMalloc when none is given, or a
// caller-owned region (the serving loop hands every admitted job a slice of
// one pre-allocated arena, so no allocation happens while other jobs run)
class DeviceAllocator {
public:
    virtual ~DeviceAllocator() = default;
    virtual void *allocate(size_t bytes) = 0;  // throws when it cannot
    virtual void release(void *p) = 0;
};

// caller-owned device arrays a runner uses in place of its own buffers (all
// or none): the director (caller layout), the mask as float64 (>= 0 is LC,
// may be null), the interface weights (may be null) and the output E, which
// also holds the input fields of "Beam profile": "None"
struct DeviceIO {
    const double *lc = nullptr;
    const double *mask = nullptr;
    const double *weights = nullptr;
    c64 *E = nullptr;
};

struct Options {
    int precision = 64;  // 64: complex128 march; 32: complex64 march, FP64 coefficients
    int wl_batch = 1;    // wavelengths marched together (<= 0: all)
    int use_graph = 0;   // capture the z-march as one hipGraph and replay it
    int device = -1;     // HIP device (-1: current)
    int profile = 0;     // time every kernel family with events (no graph)
    Tuning tuning;       // x_chunks == 0: default shapes
    DeviceAllocator *allocator = nullptr;  // nullptr: hipMalloc / hipFree
    DeviceIO dev;                // dev.lc set: inputs and output are the caller's device arrays
    bool no_timing = false;      // no HIP events (stage times stay 0)
    bool no_side_stream = false; // no second stream (coefficients on the march stream)
    int coef_batch = 0;          // slabs per coefficient launch on the march stream (0: by size)
};

// per-kernel-family device time of one propagate() (profile mode); the slab
// material is evaluated inside the coefficient kernel, so KF_MATERIAL stays 0
enum KernelFamily {
    KF_MATERIAL = 0, KF_COEFS, KF_BOUNDARY, KF_RESERVED, KF_EXPLICIT, KF_SOLVE_X, KF_SOLVE_Y, KF_OUTPUT, KF_FILTER,
    KF_SETUP, KF_COUNT
};

struct Stats {
    double setup_ms = 0, march_ms = 0, output_ms = 0, total_ms = 0;
    double updates = 0;
    int64_t footprint_bytes = 0;
    double family_ms[KF_COUNT] = {0};
    int64_t family_launches[KF_COUNT] = {0};
};

class Runner {
public:
    virtual ~Runner() = default;
    // caller-supplied input fields (beam profile "None"), complex128, caller layout
    virtual void upload_fields(const c64 *E_host) = 0;
    // full propagation on the device, inputs already resident
    virtual void propagate(Stats *stats) = 0;
    // output fields (caller layout) back to the host
    virtual void download(c64 *E_host) = 0;
    // device copy of the output (caller layout), valid after propagate
    virtual const c64 *device_output() const = 0;
    virtual int64_t footprint() const = 0;
    // new line-solver launch shapes (throws, keeping the old ones, when the
    // device cannot run them)
    virtual void retune(const Tuning &t) = 0;
    virtual hipStream_t stream() const = 0;
    // notice the run produced while preparing its inputs (the reference's
    // printed messages that do not stop it), empty if none
    virtual const std::string &notice() const = 0;
};

std::unique_ptr<Runner> make_runner(const Problem &p, const Options &o, hipStream_t stream = nullptr);

// upper bound of the device bytes a runner of this problem allocates (every
// buffer rounded up to kAllocAlign); used for serving admission
constexpr size_t kAllocAlign = 256;
int64_t estimate_footprint(const Problem &p, const Options &o);

// interface weights of p's compiled mask formula (p.weights_on_device) and
// LC flags (p.mask) computed on the current device, downloaded into w (left
// empty, with the reference's notice in *warning, on a mismatch)
void device_interface_weights(const Problem &p, std::vector<double> &w, std::string *warning);

}  // namespace nmk
