#include "server.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>

namespace nmk {

namespace {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at " + what);
}

// bump allocator inside the arena slice of one admitted job; everything is
// returned at once when the job ends
class SliceAllocator final : public DeviceAllocator {
public:
    SliceAllocator(char *base, size_t size) : base_(base), size_(size) {}
    void *allocate(size_t bytes) override {
        const size_t at = (used_ + kAllocAlign - 1) / kAllocAlign * kAllocAlign;
        if (at + bytes > size_)
            throw std::runtime_error("serving: job exceeded its admitted device-memory footprint");
        used_ = at + bytes;
        return base_ + at;
    }
    void release(void *) override {}

private:
    char *base_;
    size_t size_, used_ = 0;
};

// bytes reserved per job beyond the footprint estimate
constexpr size_t kJobSlack = 64 * kAllocAlign;

}  // namespace

// ---------------------------------------------------------------- ranges
RangeAllocator::RangeAllocator(size_t size) : size_(size) {
    if (size) free_[0] = size;
}

bool RangeAllocator::allocate(size_t bytes, size_t *offset) {
    bytes = (bytes + kAllocAlign - 1) / kAllocAlign * kAllocAlign;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        taken_[off] = bytes;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        *offset = off;
        return true;
    }
    return false;
}

void RangeAllocator::release(size_t offset) {
    auto t = taken_.find(offset);
    if (t == taken_.end()) return;
    size_t off = t->first, len = t->second;
    taken_.erase(t);
    used_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
        len += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += len;
            return;
        }
    }
    free_[off] = len;
}

// ---------------------------------------------------------------- server
Server::Server(const ServerConfig &cfg) : cfg_(cfg), t0_(std::chrono::steady_clock::now()) {
    if (cfg_.max_active < 1) throw BpmError("serving: max_active must be at least 1");
    if (cfg_.max_bypass < 0) cfg_.max_bypass = 0;
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (n <= 0) throw std::runtime_error("serving: no HIP device");
    if (cfg_.device < 0) hip_check(hipGetDevice(&cfg_.device), "hipGetDevice");
    hip_check(hipSetDevice(cfg_.device), "hipSetDevice");
    size_t freeb = 0, totalb = 0;
    hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
    size_t budget = cfg_.hbm_budget > 0 ? (size_t)cfg_.hbm_budget : (size_t)(0.8 * (double)freeb);
    budget = budget / kAllocAlign * kAllocAlign;
    if (budget == 0 || budget > freeb) throw BpmError("serving: the HBM budget exceeds the free device memory");
    hip_check(hipMalloc((void **)&arena_, budget), "hipMalloc(arena)");
    ranges_.reset(new RangeAllocator(budget));
    st_.arena_bytes = (int64_t)budget;
    for (int w = 0; w < cfg_.max_active; w++) {
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        streams_.push_back(s);
    }
    for (int w = 0; w < cfg_.max_active; w++) workers_.emplace_back(&Server::worker, this, w);
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        for (auto &j : queue_) {
            j->info.state = JOB_CANCELLED;
            j->info.t_end_ms = now_ms();
            st_.cancelled++;
        }
        queue_.clear();
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    for (auto &t : workers_) t.join();
    for (auto s : streams_) (void)hipStreamDestroy(s);
    if (arena_) (void)hipFree(arena_);
}

double Server::now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
}

int64_t Server::submit(const std::string &json, const char *mask_formula, const double *lc, int64_t n_lc,
                       const double *mask, int64_t n_mask, c64 *E, int64_t n_E, const Options &opt) {
    auto job = std::make_shared<Job>();
    job->info.t_submit_ms = now_ms();
    if (!lc || n_lc < 0 || !E || n_E < 0) throw BpmError("serving: missing input or output array");
    const Settings s = parse_settings(json);
    job->problem = build_problem(s, lc, n_lc, mask, mask ? n_mask : 0, mask ? mask_formula : nullptr, n_E, nullptr,
                                 /*defer_mask=*/true);
    job->opt = opt;
    job->opt.device = -1;
    job->opt.profile = 0;
    job->opt.allocator = nullptr;
    job->E = E;
    const int64_t need = estimate_footprint(job->problem, job->opt) + (int64_t)kJobSlack;
    job->info.footprint = need;
    job->info.updates = (double)job->problem.Nx * job->problem.Ny * (job->problem.Nz - 1) *
                        (double)job->problem.wl.size() * (double)job->problem.qx.size() * 2;
    if ((size_t)need > ranges_->size())
        throw BpmError("serving: the job needs " + std::to_string(need) + " bytes of device memory, the arena holds " +
                       std::to_string(ranges_->size()));
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw BpmError("serving: the server is shutting down");
    job->id = next_id_++;
    jobs_[job->id] = job;
    queue_.push_back(job);
    st_.submitted++;
    schedule_locked();
    return job->id;
}

// Admits queued jobs in arrival order while slots and a contiguous arena
// range are free.  A job that does not fit is passed by later jobs that do,
// until it has been passed max_bypass times; then admission waits for it.
void Server::schedule_locked() {
    std::vector<Job *> skipped;
    for (auto it = queue_.begin(); it != queue_.end() && active_ < cfg_.max_active && !stop_;) {
        Job &j = **it;
        size_t off = 0;
        if (ranges_->allocate((size_t)j.info.footprint, &off)) {
            j.region = off;
            j.info.state = JOB_RUNNING;
            j.info.t_admit_ms = now_ms();
            for (Job *s : skipped) s->bypassed++;
            ready_.push_back(*it);
            it = queue_.erase(it);
            active_++;
            st_.peak_active = std::max(st_.peak_active, active_);
            continue;
        }
        if (j.bypassed >= cfg_.max_bypass) break;
        skipped.push_back(&j);
        ++it;
    }
    st_.arena_used = (int64_t)ranges_->used();
    st_.arena_peak = (int64_t)ranges_->peak();
    if (!ready_.empty()) cv_work_.notify_all();
}

void Server::run_job(Job &job, hipStream_t st) {
    // host preparation of the job: per-voxel mask and interface weights
    std::string notice;
    prepare_mask(job.problem, &notice);
    SliceAllocator slice(arena_ + job.region, (size_t)job.info.footprint);
    Options o = job.opt;
    o.allocator = &slice;
    std::unique_ptr<Runner> r = make_runner(job.problem, o, st);
    if (notice.empty()) notice = r->notice();
    // inputs are on the device now; the host copies are no longer needed
    job.problem.lc = nullptr;
    std::vector<uint8_t>().swap(job.problem.mask);
    std::vector<double>().swap(job.problem.weights);
    if (job.problem.beam == BeamKind::None) r->upload_fields(job.E);
    {
        std::lock_guard<std::mutex> lk(mu_);
        job.info.t_ready_ms = now_ms();
        job.info.notice = notice;
    }
    Stats s;
    r->propagate(&s);
    r->download(job.E);
    job.info.device_ms = s.total_ms;
}

void Server::worker(int w) {
    (void)hipSetDevice(cfg_.device);
    for (;;) {
        std::shared_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return !ready_.empty() || (stop_ && active_ == 0); });
            if (ready_.empty()) return;
            job = ready_.front();
            ready_.pop_front();
        }
        int state = JOB_DONE, status = 0;
        std::string error;
        try {
            run_job(*job, streams_[w]);
        } catch (const BpmError &e) {
            state = JOB_FAILED;
            status = 1;
            error = e.what();
        } catch (const std::bad_alloc &) {
            state = JOB_FAILED;
            status = 3;
            error = "serving: host memory exhausted";
        } catch (const std::exception &e) {
            state = JOB_FAILED;
            status = std::string(e.what()).rfind("HIP error", 0) == 0 ? 2 : 4;
            error = e.what();
        }
        // a job that failed part-way may have kernels or copies still queued
        // on its stream that write into its arena range: the range is handed
        // to the next job only once that stream has drained
        const hipError_t drained = hipStreamSynchronize(streams_[w]);
        if (drained != hipSuccess && state == JOB_DONE) {
            state = JOB_FAILED;
            status = 2;
            error = std::string("HIP error: ") + hipGetErrorString(drained) + " at hipStreamSynchronize";
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            ranges_->release(job->region);
            active_--;
            job->info.state = state;
            job->info.status = status;
            job->info.error = error;
            job->info.t_end_ms = now_ms();
            job->problem = Problem();
            if (state == JOB_DONE) st_.completed++;
            else st_.failed++;
            schedule_locked();
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

int Server::wait(int64_t id, double timeout_ms, JobInfo *info) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return -1;
    std::shared_ptr<Job> job = it->second;
    auto finished = [&] { return job->info.state >= JOB_DONE; };
    if (timeout_ms < 0) {
        cv_done_.wait(lk, finished);
    } else if (!cv_done_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
        if (info) *info = job->info;
        return 1;
    }
    if (info) *info = job->info;
    return 0;
}

bool Server::cancel(int64_t id, bool *known) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    *known = it != jobs_.end();
    if (!*known) return false;
    Job &j = *it->second;
    if (j.info.state != JOB_QUEUED) return false;
    queue_.erase(std::remove(queue_.begin(), queue_.end(), it->second), queue_.end());
    j.info.state = JOB_CANCELLED;
    j.info.t_end_ms = now_ms();
    j.problem = Problem();
    st_.cancelled++;
    schedule_locked();
    lk.unlock();
    cv_done_.notify_all();
    return true;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ServerStats s = st_;
    s.running = active_;
    s.queued = (int64_t)queue_.size();
    s.arena_used = (int64_t)ranges_->used();
    s.arena_peak = (int64_t)ranges_->peak();
    return s;
}

}  // namespace nmk
