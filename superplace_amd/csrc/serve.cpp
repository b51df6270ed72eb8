// Many placement jobs on one GPU: a scheduler that admits jobs by device
// footprint, keeps every admitted job's design and optimizer state resident,
// and advances all of them together.  No counterpart in the reference (it
// places one design per process).
//
// Each admitted job is one placement instance (EngineBase).  FP32 jobs on
// power-of-two bin grids are batched across iterations: they advance
// together in launches of one captured graph whose kernels each serve all of
// them (BatchSet), two launches in flight.  When a batched job's stage ends,
// or set-ups have finished in the meantime, the scheduler lets the launch in
// flight end, retires the finished jobs, and rebuilds the batch with the
// newcomers, so that a job arriving while others run joins them at the next
// launch boundary.  Other jobs (FP64, grids that are not powers of two) run
// on their own streams, each driving its device-side loop in launches of its
// own graph, kAsyncDepth in flight.  Jobs are admitted as soon as a slot and
// enough of the device budget are free; the host set-ups of the next jobs
// (design upload, initial evaluations) run on worker threads meanwhile, up
// to kSetupThreads at once.  Each job's result is the same as when it runs
// alone: an instance's kernels touch only its own buffers and every kernel
// is deterministic.
#include <algorithm>
#include <chrono>
#include <future>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "abi.h"
#include "engine.h"

namespace sphip {
namespace {

// A job's own copy of the caller's design arrays.
struct DesignCopy {
    sp_hip_design view{};
    std::vector<float> x, y, w, h, offx, offy, rowLy, rowH, rowLx, rowHx;
    std::vector<uint8_t> flags;
    std::vector<int32_t> pin2node, pin2net, netStart, netPins, nodeStart, nodePins;

    explicit DesignCopy(const sp_hip_design& d) {
        validateDesign(d);
        const size_t total = static_cast<size_t>(d.num_nodes) + d.num_fillers;
        const size_t P = d.num_pins, E = d.num_nets, R = d.num_rows;
        auto take = [](auto& v, const auto* src, size_t n) {
            v.assign(src, src + n);
            return v.data();
        };
        view = d;
        view.node_x = take(x, d.node_x, total);
        view.node_y = take(y, d.node_y, total);
        view.node_w = take(w, d.node_w, total);
        view.node_h = take(h, d.node_h, total);
        view.node_flags = take(flags, d.node_flags, total);
        view.pin_offset_x = P ? take(offx, d.pin_offset_x, P) : nullptr;
        view.pin_offset_y = P ? take(offy, d.pin_offset_y, P) : nullptr;
        view.pin2node = P ? take(pin2node, d.pin2node, P) : nullptr;
        view.pin2net = P ? take(pin2net, d.pin2net, P) : nullptr;
        view.net2pin_start = take(netStart, d.net2pin_start, E + 1);
        view.flat_net2pin = P ? take(netPins, d.flat_net2pin, P) : nullptr;
        view.node2pin_start = take(nodeStart, d.node2pin_start, total + 1);
        view.flat_node2pin = P ? take(nodePins, d.flat_node2pin, P) : nullptr;
        if (R > 0 && (!d.row_ly || !d.row_height || !d.row_lx || !d.row_hx))
            throw InvalidArg("design: missing row arrays");
        view.row_ly = R ? take(rowLy, d.row_ly, R) : nullptr;
        view.row_height = R ? take(rowH, d.row_height, R) : nullptr;
        view.row_lx = R ? take(rowLx, d.row_lx, R) : nullptr;
        view.row_hx = R ? take(rowHx, d.row_hx, R) : nullptr;
    }
};

struct Job {
    int id = 0;
    std::unique_ptr<DesignCopy> design;
    sp_hip_gp_config cfg{};
    sp_hip_job_info info{};
    sp_hip_gp_result res{};
    std::vector<float> x, y;
    std::unique_ptr<EngineBase> eng;
};

constexpr int kSetupThreads = 4;   // job set-ups in flight at once
constexpr int kBatchSlots = 8;     // iterations per batched launch

// What a set-up thread hands back.
struct Admission {
    std::unique_ptr<EngineBase> eng;
    sp_hip_gp_result res{};  // written by the set-up thread, copied to the job once joined
    bool finished = false;   // the stage ended before its loop
    int status = SP_HIP_OK;
    std::string error;
};

class Server {
public:
    Server(int64_t budget, int maxActive)
        : budget_(budget), maxActive_(maxActive), t0_(std::chrono::steady_clock::now()) {}

    ~Server() {
        for (auto& a : admitting_) a.second.wait();
    }

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    int submit(const sp_hip_design& d, const sp_hip_gp_config& cfg, double arrival) {
        if (cfg.loop_mode == 0) throw InvalidArg("a served job needs loop_mode 1 or 2");
        if (!std::isfinite(arrival)) throw InvalidArg("arrival time must be finite");
        auto job = std::make_unique<Job>();
        job->id = static_cast<int>(jobs_.size());
        job->design = std::make_unique<DesignCopy>(d);
        job->cfg = cfg;
        job->info.state = SP_HIP_JOB_QUEUED;
        job->info.status = SP_HIP_OK;
        job->info.device_bytes = static_cast<int64_t>(footprintEstimate(job->design->view, cfg));
        job->info.arrival_s = arrival;
        job->info.admit_s = -1.0;
        job->info.finish_s = -1.0;
        job->res.rollback_iter = -1;
        queued_.push_back(job->id);
        jobs_.push_back(std::move(job));
        return jobs_.back()->id;
    }

    // One scheduling round; returns whether anything changed.
    bool pump() {
        bool progress = false;
        // finished set-ups join the running jobs
        for (size_t k = 0; k < admitting_.size();) {
            if (admitting_[k].second.wait_for(std::chrono::seconds(0)) != std::future_status::ready) {
                ++k;
                continue;
            }
            Admission a = admitting_[k].second.get();
            Job& j = *jobs_[admitting_[k].first];
            j.res = a.res;
            admitting_.erase(admitting_.begin() + static_cast<long>(k));
            progress = true;
            if (a.status != SP_HIP_OK) {
                fail(j, a.status, a.error);
            } else {
                used_ += static_cast<int64_t>(a.eng->deviceBytes()) - j.info.device_bytes;
                j.info.device_bytes = static_cast<int64_t>(a.eng->deviceBytes());
                j.eng = std::move(a.eng);
                if (a.finished) {
                    complete(j, false);
                } else if (j.eng->batchable()) {
                    j.info.state = SP_HIP_JOB_RUNNING;
                    joining_.push_back(j.id);
                } else {
                    j.info.state = SP_HIP_JOB_RUNNING;
                    running_.push_back(j.id);
                    topUp(j);
                }
            }
        }
        progress = pumpBatch() || progress;
        // top up the running jobs' launches, retire the finished ones
        for (size_t k = 0; k < running_.size();) {
            Job& j = *jobs_[running_[k]];
            int st = 0;
            try {
                st = advance(j);
            } catch (...) {
                st = -2;
                failCurrent(j);
            }
            if (st != 0) {
                running_.erase(running_.begin() + static_cast<long>(k));
                progress = true;
            } else {
                ++k;
            }
        }
        // start the set-up of the next jobs that have arrived, while they fit
        // (up to kSetupThreads at once, each on a thread of its own)
        const double t = now();
        std::stable_sort(queued_.begin(), queued_.end(), [&](int a, int b) {
            return jobs_[a]->info.arrival_s < jobs_[b]->info.arrival_s;
        });
        while (!queued_.empty() && static_cast<int>(admitting_.size()) < kSetupThreads) {
            Job& j = *jobs_[queued_.front()];
            if (j.info.device_bytes > budget_) {
                queued_.erase(queued_.begin());
                fail(j, SP_HIP_EOOM, "job footprint exceeds the device budget");
                progress = true;
                continue;
            }
            if (!(j.info.arrival_s <= t) ||
                static_cast<int>(running_.size() + admitting_.size() + batched_.size() +
                                 joining_.size()) >= maxActive_ ||
                used_ + j.info.device_bytes > budget_)
                break;
            queued_.erase(queued_.begin());
            used_ += j.info.device_bytes;
            j.info.admit_s = t;
            j.info.state = SP_HIP_JOB_RUNNING;
            const sp_hip_design* d = &j.design->view;
            const sp_hip_gp_config cfg = j.cfg;
            const sp_hip_gp_result start = j.res;
            admitting_.emplace_back(j.id, std::async(std::launch::async, [d, cfg, start] {
                Admission a;
                a.res = start;
                a.status = guarded([&] {
                    a.eng = makeEngine(*d, cfg);
                    a.finished = a.eng->begin(&a.res);
                });
                if (a.status != SP_HIP_OK) {
                    a.error = g_lastError;
                    a.eng.reset();
                }
                return a;
            }));
            progress = true;
        }
        return progress;
    }

    void counts(int32_t* active, int32_t* queued, int32_t* finished) const {
        if (active)
            *active = static_cast<int32_t>(running_.size() + admitting_.size() + batched_.size() +
                                           joining_.size());
        if (queued) *queued = static_cast<int32_t>(queued_.size());
        if (finished) *finished = finished_;
    }

    // Pumps until every submitted job has finished or failed.
    void drain(double timeout) {
        const double end = now() + timeout;
        while (finished_ < static_cast<int>(jobs_.size())) {
            if (now() > end) throw InvalidArg("drain: time limit reached");
            if (pump()) continue;
            if (running_.empty() && admitting_.empty() && batched_.empty() && joining_.empty() &&
                !queued_.empty()) {
                // idle until the next arrival
                double next = end;
                for (int id : queued_) next = std::min(next, jobs_[id]->info.arrival_s);
                const double wait = next - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            } else {
                std::this_thread::yield();
            }
        }
    }

    const Job& job(int id) const {
        if (id < 0 || id >= static_cast<int>(jobs_.size())) throw InvalidArg("no such job");
        return *jobs_[id];
    }

    void release(int id) {
        if (id < 0 || id >= static_cast<int>(jobs_.size())) throw InvalidArg("no such job");
        Job& j = *jobs_[id];
        if (j.info.state != SP_HIP_JOB_DONE && j.info.state != SP_HIP_JOB_FAILED)
            throw InvalidArg("job still queued or running");
        j.design.reset();
        std::vector<float>().swap(j.x);
        std::vector<float>().swap(j.y);
    }

private:
    // The batched jobs: reads the oldest launch's done flags once it has
    // ended; on a change of members (stages ended, set-ups joined) lets the
    // other launch end, retires the finished jobs and rebuilds the batch;
    // keeps kAsyncDepth launches in flight.  Returns whether anything changed.
    bool pumpBatch() {
        if (batched_.empty() && joining_.empty()) return false;
        bool progress = false;
        try {
            if (!batch_) batch_ = std::make_unique<BatchSet>(kBatchSlots);
            std::vector<int> ended(batched_.size(), 0);   // 1 done, 2 over its limits
            auto merge = [&] {
                for (size_t q = 0; q < ended.size(); ++q)
                    if (state_[q] != 0 && ended[q] == 0) ended[q] = state_[q];
            };
            if (batch_->inFlight() > 0) {
                if (batch_->poll(false, &state_) < 0) return false;
                merge();
                progress = true;
            }
            const bool change =
                !joining_.empty() || std::any_of(ended.begin(), ended.end(), [](int e) { return e; });
            if (change) {
                while (batch_->inFlight() > 0) {
                    batch_->poll(true, &state_);
                    merge();
                }
                std::vector<int> keep;
                for (size_t q = 0; q < batched_.size(); ++q) {
                    Job& j = *jobs_[batched_[q]];
                    if (ended[q] == 1) {
                        try {
                            complete(j, true);
                        } catch (...) {
                            failCurrent(j);
                        }
                    } else if (ended[q] == 2) {
                        fail(j, SP_HIP_EINTERNAL, "batched loop did not terminate");
                    } else {
                        keep.push_back(batched_[q]);
                    }
                }
                keep.insert(keep.end(), joining_.begin(), joining_.end());
                joining_.clear();
                batched_ = keep;
                std::vector<EngineBase*> members;
                for (int id : batched_) members.push_back(jobs_[id]->eng.get());
                batch_->assign(members);
                progress = true;
            }
            while (batch_->size() > 0 && batch_->inFlight() < kAsyncDepth) {
                batch_->launch();
                progress = true;
            }
        } catch (...) {
            // a device or set-up error leaves the batch unusable: its jobs fail
            std::vector<int> all(batched_);
            all.insert(all.end(), joining_.begin(), joining_.end());
            batched_.clear();
            joining_.clear();
            batch_.reset();
            for (int id : all) {
                Job& j = *jobs_[id];
                if (j.info.state == SP_HIP_JOB_RUNNING) fail(j, SP_HIP_EDEVICE, "batched launch failed");
            }
            progress = true;
        }
        return progress;
    }

    void topUp(Job& j) {
        while (j.eng->inFlight() < kAsyncDepth) j.eng->launchAsync();
    }

    // 0: still running; 1: completed now.
    int advance(Job& j) {
        for (;;) {
            const int st = j.eng->pollAsync(false);
            if (st < 0) return 0;
            if (st == 1) {
                complete(j, true);
                return 1;
            }
            j.eng->launchAsync();
        }
    }

    void complete(Job& j, bool loop) {
        if (loop) j.eng->finishAsync(&j.res);
        const size_t total = j.design->x.size();
        j.x.resize(total);
        j.y.resize(total);
        j.eng->positions(j.x.data(), j.y.data());
        retire(j, SP_HIP_JOB_DONE, SP_HIP_OK);
    }

    void failCurrent(Job& j) {
        int status = SP_HIP_EINTERNAL;
        try {
            throw;
        } catch (const InvalidArg&) {
            status = SP_HIP_EINVAL;
        } catch (const NonFinite&) {
            status = SP_HIP_ENONFINITE;
        } catch (const HipError&) {
            status = SP_HIP_EDEVICE;
        } catch (...) {
        }
        fail(j, status, "job failed");
    }

    void fail(Job& j, int status, const std::string& why) {
        (void)why;
        retire(j, SP_HIP_JOB_FAILED, status);
    }

    void retire(Job& j, int state, int status) {
        if (j.eng) {
            used_ -= j.info.device_bytes;
            j.eng.reset();   // frees the instance's device memory
        } else if (j.info.admit_s >= 0 && state == SP_HIP_JOB_FAILED) {
            used_ -= j.info.device_bytes;   // reserved at admission, never built
        }
        j.info.state = state;
        j.info.status = status;
        j.info.finish_s = now();
        ++finished_;
    }

    int64_t budget_;
    int maxActive_;
    std::chrono::steady_clock::time_point t0_;
    std::vector<std::unique_ptr<Job>> jobs_;
    std::vector<int> queued_, running_;
    std::vector<int> batched_, joining_;   // members of batch_ (in its order), waiting to join
    std::unique_ptr<BatchSet> batch_;
    std::vector<int> state_;
    std::vector<std::pair<int, std::future<Admission>>> admitting_;   // (job, set-up)
    int64_t used_ = 0;
    int finished_ = 0;
};

}  // namespace
}  // namespace sphip

struct sp_hip_server {
    std::unique_ptr<sphip::Server> impl;
};

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_hip_gp_footprint(const sp_hip_design* d, const sp_hip_gp_config* cfg, int64_t* bytes) {
    return guarded([&] {
        if (!d || !cfg || !bytes) throw InvalidArg("null argument");
        *bytes = static_cast<int64_t>(sphip::footprintEstimate(*d, *cfg));
    });
}

int sp_hip_server_create(int64_t hbm_budget, int32_t max_active, sp_hip_server** out) {
    return guarded([&] {
        if (!out) throw InvalidArg("null argument");
        *out = nullptr;
        if (hbm_budget <= 0 || max_active <= 0) throw InvalidArg("budget and slots must be positive");
        sphip::requireDevice();
        auto s = std::make_unique<sp_hip_server>();
        s->impl = std::make_unique<sphip::Server>(hbm_budget, max_active);
        *out = s.release();
    });
}

int sp_hip_server_submit(sp_hip_server* s, const sp_hip_design* d, const sp_hip_gp_config* cfg,
                         double arrival_s, int32_t* job) {
    return guarded([&] {
        if (!s || !d || !cfg || !job) throw InvalidArg("null argument");
        *job = s->impl->submit(*d, *cfg, arrival_s);
    });
}

int sp_hip_server_pump(sp_hip_server* s, double* now_s, int32_t* active, int32_t* queued,
                       int32_t* finished) {
    return guarded([&] {
        if (!s) throw InvalidArg("null argument");
        s->impl->pump();
        s->impl->counts(active, queued, finished);
        if (now_s) *now_s = s->impl->now();
    });
}

int sp_hip_server_drain(sp_hip_server* s, double timeout_s) {
    return guarded([&] {
        if (!s) throw InvalidArg("null argument");
        s->impl->drain(timeout_s);
    });
}

int sp_hip_server_now(const sp_hip_server* s, double* now_s) {
    return guarded([&] {
        if (!s || !now_s) throw InvalidArg("null argument");
        *now_s = s->impl->now();
    });
}

int sp_hip_server_job(const sp_hip_server* s, int32_t job, sp_hip_job_info* info,
                      sp_hip_gp_result* res, float* x, float* y) {
    return guarded([&] {
        if (!s) throw InvalidArg("null argument");
        const auto& j = s->impl->job(job);
        if (info) *info = j.info;
        if (res) *res = j.res;
        if ((x || y) && j.info.state == SP_HIP_JOB_DONE) {
            if (j.x.empty() && !j.design) throw InvalidArg("job result released");
            if (x) std::copy(j.x.begin(), j.x.end(), x);
            if (y) std::copy(j.y.begin(), j.y.end(), y);
        }
    });
}

int sp_hip_server_release(sp_hip_server* s, int32_t job) {
    return guarded([&] {
        if (!s) throw InvalidArg("null argument");
        s->impl->release(job);
    });
}

void sp_hip_server_destroy(sp_hip_server* s) { delete s; }

}  // extern "C"
