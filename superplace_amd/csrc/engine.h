// The placement instance behind sp_hip_gp, as seen by the other host code
// of the library (the scheduler of concurrent instances).
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "../../include/superplace_hip.h"

namespace sphip {

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual void info(sp_hip_gp_info* out) const = 0;
    virtual void optIndex(int32_t* out) const = 0;
    virtual void evalHost(const float* x, const float* y, float lambda, float gamma,
                          sp_hip_eval_scalars* out, sp_hip_eval_debug* dbg) = 0;
    virtual void run(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows) = 0;
    virtual void positions(float* x, float* y) const = 0;
    virtual void setProfiling(int mask) = 0;
    virtual void kernelStats(double* ms, int64_t* launches, double* bytes, double* flops) const = 0;
    // The stage loop in three parts, for a caller that interleaves several
    // instances (device-driven loop modes only): begin() runs the host part
    // before the loop and returns true when the stage ended there; then
    // launchAsync() enqueues one launch of the loop (at most kAsyncDepth in
    // flight) and pollAsync() reports the oldest one (-1 still running, 0
    // finished with the stage still going, 1 stage done); finishAsync()
    // completes the result once the stage is done.
    virtual bool begin(sp_hip_gp_result* res) = 0;
    virtual void launchAsync() = 0;
    virtual int pollAsync(bool wait) = 0;
    virtual int inFlight() const = 0;
    virtual void finishAsync(sp_hip_gp_result* res) = 0;
    virtual size_t deviceBytes() const = 0;
    // Whether the instance can join batched launches (FP32, power-of-two
    // bin grid).
    virtual bool batchable() const = 0;
};

constexpr int kAsyncDepth = 2;   // loop launches in flight per instance

// Advances many FP32 instances (each created with a device-driven
// loop_mode, power-of-two bin grids, begin() already called: looping[i] is
// whether instance i's stage still has its loop to run) through their
// stages with batched launches: one hipGraph of `slots` slots, each kernel
// launched once for all of them.  res[i] receives instance i's result,
// equal to its own run's.
void runBatch(const std::vector<EngineBase*>& engines, const std::vector<char>& looping,
              sp_hip_gp_result* res, int slots);

// Batched launches over a set of FP32 instances that may change between
// launches (continuous batching).  Every launch runs `slots` slots, each
// kernel of a slot launched once for all the members (the FFT passes once per
// grid size); a member whose stage has ended is skipped by every kernel until
// it leaves.  Each member's result equals its own run's: its kernels touch
// only its own buffers and are deterministic, and every member starts at a
// launch boundary as its own loop does.
class BatchSet {
public:
    explicit BatchSet(int slots);
    ~BatchSet();
    BatchSet(const BatchSet&) = delete;
    BatchSet& operator=(const BatchSet&) = delete;

    // The members from now on (each begun, with its stage loop still to
    // run).  Waits for the launches in flight, then rebuilds the block maps
    // and the captured graph.
    void assign(const std::vector<EngineBase*>& members);
    // Enqueues one launch for all members (at most kAsyncDepth in flight).
    void launch();
    // The oldest launch in flight: -1 while it runs (wait = false), else 0
    // with state[j] for member j: 0 still running, 1 stage done, 2 not done
    // after more launches than its iteration and backtrack limits allow.
    int poll(bool wait, std::vector<int>* state);
    int inFlight() const;
    int size() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// A new instance on its own stream (precision from cfg.precision).
std::unique_ptr<EngineBase> makeEngine(const sp_hip_design& d, const sp_hip_gp_config& cfg);

// Device bytes an instance of this design and configuration allocates, an
// upper bound of EngineBase::deviceBytes() once its stage loop has run.
size_t footprintEstimate(const sp_hip_design& d, const sp_hip_gp_config& cfg);

}  // namespace sphip
