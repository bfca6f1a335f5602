// Device-resident serving loop of the bpm-solver propagation step.
//
// The reference runs one blocking propagation per call
// (PythonBackend/bpm_backend.cpp:7-164); this server keeps a queue of
// micrograph jobs, admits them onto one GPU by their device-memory footprint
// as memory and execution slots free up (continuous batching), and runs the
// admitted jobs concurrently, each on its own HIP stream with its state
// resident in a slice of one pre-allocated HBM arena.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "settings.h"

namespace nmk {

struct ServerConfig {
    int device = -1;          // HIP device (-1: current)
    int max_active = 4;       // jobs on the device at once (one host worker and stream each)
    int64_t hbm_budget = 0;   // bytes of the arena; 0: 80 % of the free device memory at start
    int max_bypass = 8;       // times a queued job may be overtaken by later jobs that fit
};

enum JobState : int { JOB_QUEUED = 0, JOB_RUNNING = 1, JOB_DONE = 2, JOB_FAILED = 3, JOB_CANCELLED = 4 };

struct JobInfo {
    int state = JOB_QUEUED;
    int status = 0;            // nmk_bpm_status of a failed job
    double t_submit_ms = 0;    // host steady clock, ms since the server started
    double t_admit_ms = -1;
    double t_ready_ms = -1;    // host preparation done and inputs on the device
    double t_end_ms = -1;
    double device_ms = 0;      // device time of the propagation
    double updates = 0;        // voxel-slice updates of the job
    int64_t footprint = 0;     // admitted bytes
    std::string error;
    std::string notice;        // what the reference prints on success (mask formula mismatch)
};

struct ServerStats {
    int64_t submitted = 0, completed = 0, failed = 0, cancelled = 0, running = 0, queued = 0;
    int64_t arena_bytes = 0, arena_used = 0, arena_peak = 0;
    int peak_active = 0;
};

// first-fit allocator of byte ranges [0, size) with coalescing of freed ranges
class RangeAllocator {
public:
    explicit RangeAllocator(size_t size);
    bool allocate(size_t bytes, size_t *offset);
    void release(size_t offset);
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    size_t size_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    std::map<size_t, size_t> taken_;  // offset -> length
};

class Server {
public:
    explicit Server(const ServerConfig &cfg);
    ~Server();  // cancels queued jobs and waits for the running ones
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Validates the job (settings, array sizes, the mask formula's syntax,
    // footprint against the arena) and queues it; throws BpmError on invalid
    // input or a job that can never fit.  Nothing of size O(volume) happens
    // here: the worker that runs the job builds its mask and interface
    // weights and uploads lc when the job is admitted.  lc, mask and E
    // (complex128, caller layout) must therefore stay valid until the job has
    // finished: E is read first with "Beam profile": "None" and receives the
    // output fields.
    int64_t submit(const std::string &json, const char *mask_formula, const double *lc, int64_t n_lc,
                   const double *mask, int64_t n_mask, c64 *E, int64_t n_E, const Options &opt);
    // 0: the job has finished (done, failed or cancelled); 1: timed out
    // (timeout_ms < 0 waits without limit); -1: unknown id
    int wait(int64_t id, double timeout_ms, JobInfo *info);
    // cancels a job that is still queued (true), false once it has been admitted
    bool cancel(int64_t id, bool *known);
    ServerStats stats();
    double now_ms() const;

private:
    struct Job {
        int64_t id = 0;
        Problem problem;  // lc and mask borrowed from the caller until the job ends
        Options opt;
        c64 *E = nullptr;
        size_t region = 0;
        int bypassed = 0;
        JobInfo info;
    };
    void schedule_locked();
    void worker(int w);
    void run_job(Job &job, hipStream_t st);

    ServerConfig cfg_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Job>> queue_, ready_;
    std::map<int64_t, std::shared_ptr<Job>> jobs_;
    std::vector<std::thread> workers_;
    std::vector<hipStream_t> streams_;
    int64_t next_id_ = 0;
    int active_ = 0;
    bool stop_ = false;
    char *arena_ = nullptr;
    std::unique_ptr<RangeAllocator> ranges_;
    ServerStats st_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace nmk
