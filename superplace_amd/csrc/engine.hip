// Host side of the HIP backend: the device-resident placement instance, the
// ePlace stage loop driving the kernels, the Poisson backend and the C ABI.
//
// The loop follows EPlace::run (src/gp/eplace.cpp:223-526) and
// NesterovOptimizer (src/numeric/cpu/nesterov.cpp:51-240).  Each evaluation
// is a fixed chain of kernels on the instance's stream; the host reads back
// only the ten FP64 scalars of the evaluation to take the data-dependent
// decisions (backtracking, step size, schedules, stopping).
#include <algorithm>
#include <array>
#include <atomic>
#include <exception>
#include <thread>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/superplace_hip.h"
#include "abi.h"
#include "engine.h"
#include "wa_device.h"
#include "density_device.h"
#include "gp_kernels.h"
#include "batch.h"
#include "poisson.h"
#include "devmem.h"

namespace sphip {
namespace {

template <typename T>
T* dalloc(size_t n) {
    void* p = nullptr;
    SPHIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    g_devBytes += std::max<size_t>(n, 1) * sizeof(T);
    return static_cast<T*>(p);
}

// computeGamma, src/gp/wa_wirelength.cpp:18-29, with the power taken as 10^e
// in FP64 rounded once to FP32 (the correctly rounded value), as the device
// loop's gammaOf takes it, so that host and device decisions agree bit for
// bit.  (The reference's powf rounds differently for about 0.07% of
// exponents.)
float computeGamma(float tau, float binStep) {
    float gamma = 8.0f * binStep;
    if (tau > 1.0f) {
        gamma *= 10.0f;
    } else if (tau < 0.1f) {
        gamma *= 0.1f;
    } else {
        const float e = (20.0f / 9.0f) * (tau - 0.1f) - 1.0f;
        gamma *= static_cast<float>(std::pow(10.0, static_cast<double>(e)));
    }
    return std::max(gamma, 1e-4f);
}

// lambdaMultiplier, src/gp/lambda_schedule.h:22-25.
float lambdaMultiplier(double dH, double ref) {
    const double mu = std::pow(1.1, -dH / ref + 1.0);
    return static_cast<float>(std::min(1.05, std::max(0.95, mu)));
}

template <typename T>
inline T overlapT(T alx, T aly, T ahx, T ahy, T blx, T bly, T bhx, T bhy) {
    const T dx = std::min(ahx, bhx) - std::max(alx, blx);
    const T dy = std::min(ahy, bhy) - std::max(aly, bly);
    return (dx > T(0) && dy > T(0)) ? dx * dy : T(0);
}

// Movable cell and macro areas (fixed and terminal_NI nodes excluded).
void movableAreas(const sp_hip_design& d, double* cellArea, double* macroArea) {
    double c = 0, m = 0;
    for (int i = 0; i < d.num_movable; ++i) {
        if (d.node_flags[i] & (F_FIXED | F_NI)) continue;
        const double a = static_cast<double>(d.node_w[i]) * d.node_h[i];
        if (d.node_flags[i] & F_MACRO) m += a;
        else c += a;
    }
    *cellArea = c;
    *macroArea = m;
}

// Bins per side: gp_bin_dim, or the largest power of two whose bins hold
// at least one average movable node at the target density
// (BinGrid::initialize, bin_grid.cpp:172-189).
int binDim(const sp_hip_design& d, const sp_hip_gp_config& cfg) {
    constexpr int kMaxBinDim = 1024;
    if (cfg.gp_bin_dim > 0) return std::min(cfg.gp_bin_dim, kMaxBinDim);
    double cellArea = 0, macroArea = 0;
    movableAreas(d, &cellArea, &macroArea);
    const double avgArea = d.num_movable > 0 ? (cellArea + macroArea) / d.num_movable : 1.0;
    const double idealBin = std::max(avgArea / cfg.target_density, 1e-9);
    const double coreArea = static_cast<double>(d.core_hx - d.core_lx) *
                            static_cast<double>(d.core_hy - d.core_ly);
    const double ideal = coreArea / idealBin;
    int dim = 4;
    while (dim * 2 <= kMaxBinDim && static_cast<double>(dim * 2) * (dim * 2) <= ideal) dim *= 2;
    return dim;
}

// ---------------------------------------------------------- launch shapes
// The compiled-in shapes per bin-grid size: the autotuner's last table
// (superplace_amd/shapes.json); dim 0 is the entry of sizes without one.
struct DimShapes {
    int dim;
    sp_hip_shapes shapes;
};
constexpr DimShapes kTunedShapes[] = {
    {0, {kTileChunk, kSortPeriod, 0, 8}},
    {128, {256, 4, 0, 8}},
    {256, {256, 4, 0, 8}},
    {512, {512, 2, 0, 8}},
    {1024, {512, 2, 2, 4}},
};

struct ShapeTable {
    std::mutex mu;
    std::vector<DimShapes> entries;   // overrides, then the compiled-in table
};
ShapeTable& shapeTable() {
    static ShapeTable t;
    return t;
}

void validateShapes(const sp_hip_shapes& s) {
    if (!scatterChunkOk(s.scatter_chunk))
        throw InvalidArg("shapes: scatter_chunk must be 256, 512, 1024 or 2048");
    if (s.sort_period < 1 || s.sort_period > 8) throw InvalidArg("shapes: sort_period must be 1-8");
    if (s.col_pairs != 0 && s.col_pairs != 1 && s.col_pairs != 2 && s.col_pairs != 4 &&
        s.col_pairs != 8)
        throw InvalidArg("shapes: col_pairs must be 0, 1, 2, 4 or 8");
    if (s.graph_slots < 1 || s.graph_slots > 64) throw InvalidArg("shapes: graph_slots must be 1-64");
}

// An instance's shapes: an override of its size, the compiled-in entry of
// its size, an override of size 0, the compiled-in entry of size 0.
sp_hip_shapes shapesFor(int dim) {
    std::lock_guard<std::mutex> lock(shapeTable().mu);
    const auto& over = shapeTable().entries;
    for (int want : {dim, 0}) {
        for (const auto& e : over)
            if (e.dim == want) return e.shapes;
        for (const auto& e : kTunedShapes)
            if (e.dim == want) return e.shapes;
    }
    return sp_hip_shapes{kTileChunk, kSortPeriod, 0, 8};
}

// ===========================================================================

template <typename T>
class Engine final : public EngineBase {
public:
    Engine(const sp_hip_design& d, const sp_hip_gp_config& cfg) : cfg_(cfg) {
        requireDevice();
        validateDesign(d);
        SPHIP_CHECK(hipGetDevice(&device_));
        s_ = takeStream(device_);
        const size_t before = g_devBytes;
        try {
            build(nodeOrder(d));
        } catch (...) {
            release();
            throw;
        }
        devBytes_ = g_devBytes - before;
    }
    ~Engine() override { release(); }

    void info(sp_hip_gp_info* o) const override {
        o->dim = dim_;
        o->n_opt = n_;
        o->frac_bits = fracBits_;
        o->precision = sizeof(T) == 4 ? 32 : 64;
        o->step_x = stepX_;
        o->step_y = stepY_;
        o->origin_x = originX_;
        o->origin_y = originY_;
        o->movable_area_scaled = movableAreaScaled_;
        o->device_bytes = static_cast<int64_t>(devBytes_);
        o->shapes = sh_;
        o->shapes.col_pairs = fuseWa_ ? 0 : poisson_->columnPairs();
    }

    void optIndex(int32_t* out) const override {
        for (int k = 0; k < n_; ++k) out[k] = extOf(optIdxHost_[k]);
    }

    void positions(float* x, float* y) const override {
        std::vector<T> hx(total_), hy(total_);
        downloadXY(hx.data(), hy.data());
        for (int i = 0; i < total_; ++i) {
            x[extOf(i)] = static_cast<float>(hx[i]);
            y[extOf(i)] = static_cast<float>(hy[i]);
        }
    }

    void setProfiling(int mask) override { profMask_ = mask; }
    // iterations per graph launch: the configuration's, else the shapes'
    int loopSlots() const { return cfg_.graph_slots > 0 ? cfg_.graph_slots : sh_.graph_slots; }
    void kernelStats(double* ms, int64_t* launches, double* bytes, double* flops) const override {
        for (int f = 0; f < SP_HIP_K_COUNT; ++f) {
            if (ms) ms[f] = kms_[f];
            if (launches) launches[f] = klaunch_[f];
            if (bytes) bytes[f] = kbytes_[f];
            if (flops) flops[f] = kflops_[f];
        }
    }

    // One evaluation at host positions of all nodes (EPlace scatter + objAndGrad).
    void evalHost(const float* x, const float* y, float lambda, float gamma,
                  sp_hip_eval_scalars* out, sp_hip_eval_debug* dbg) override {
        std::vector<T> hx(total_), hy(total_);
        for (int i = 0; i < total_; ++i) {
            hx[i] = x[extOf(i)];
            hy[i] = y[extOf(i)];
        }
        uploadXY(hx.data(), hy.data());
        std::vector<T> pos(2 * static_cast<size_t>(n_));
        for (int k = 0; k < n_; ++k) {
            pos[k] = hx[optIdxHost_[k]];
            pos[n_ + k] = hy[optIdxHost_[k]];
        }
        SPHIP_CHECK(hipMemcpyAsync(u_, pos.data(), pos.size() * sizeof(T), hipMemcpyHostToDevice, s_));
        curPos_ = nullptr;   // x/y now hold the raw positions, not a clamped pos buffer
        const bool wantDbg = dbg != nullptr;
        long long *dNode = nullptr, *dFill = nullptr;
        T *dForce = nullptr, *dWl = nullptr;
        if (wantDbg) {
            dNode = dalloc<long long>(nb_);
            dFill = dalloc<long long>(nb_);
            dForce = dalloc<T>(2 * static_cast<size_t>(n_));
            dWl = dalloc<T>(2 * static_cast<size_t>(n_));
        }
        EvalScalars sc;
        try {
            sc = evaluate(u_, nullptr, nullptr, g_, lambda, gamma, dNode, dFill, dForce, dWl);
            std::memcpy(out, &sc, sizeof(sc));
            if (wantDbg) {
                auto fetchQ = [&](int64_t* q, double* dens, const long long* dev, size_t cnt) {
                    if (!q && !dens) return;
                    std::vector<long long> tmp(cnt);
                    SPHIP_CHECK(hipMemcpyAsync(tmp.data(), dev, cnt * sizeof(long long),
                                               hipMemcpyDeviceToHost, s_));
                    SPHIP_CHECK(hipStreamSynchronize(s_));
                    for (size_t i = 0; i < cnt; ++i) {
                        if (q) q[i] = tmp[i];
                        if (dens) dens[i] = std::ldexp(static_cast<double>(tmp[i]), -fracBits_);
                    }
                };
                auto fetchT = [&](double* host, const T* dev, size_t cnt, long ld = -1, int rows = 0,
                                  int cols = 0) {
                    if (!host) return;
                    std::vector<T> tmp(ld > 0 ? static_cast<size_t>(rows) * ld : cnt);
                    SPHIP_CHECK(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(T),
                                               hipMemcpyDeviceToHost, s_));
                    SPHIP_CHECK(hipStreamSynchronize(s_));
                    if (ld > 0) {
                        for (int r = 0; r < rows; ++r)
                            for (int c = 0; c < cols; ++c)
                                host[static_cast<size_t>(r) * cols + c] = tmp[static_cast<size_t>(r) * ld + c];
                    } else {
                        for (size_t i = 0; i < cnt; ++i) host[i] = tmp[i];
                    }
                };
                fetchQ(dbg->node_q, dbg->node_density, dNode, nb_);
                fetchQ(dbg->filler_q, dbg->filler_density, dFill, nb_);
                fetchT(dbg->rho, poisson_->rho(), 0, poisson_->ld(), dim_, dim_);
                if (dbg->rho)
                    for (size_t b = 0; b < nb_; ++b) dbg->rho[b] += rhoShift_;
                if (dbg->field_x || dbg->field_y) {
                    std::vector<T> tmp(2 * static_cast<size_t>(poisson_->rowsPadded()) * poisson_->ld());
                    SPHIP_CHECK(hipMemcpyAsync(tmp.data(), poisson_->fieldXY(), tmp.size() * sizeof(T),
                                               hipMemcpyDeviceToHost, s_));
                    SPHIP_CHECK(hipStreamSynchronize(s_));
                    for (int r = 0; r < dim_; ++r)
                        for (int c = 0; c < dim_; ++c) {
                            const size_t at = 2 * (static_cast<size_t>(r) * poisson_->ld() + c);
                            if (dbg->field_x) dbg->field_x[static_cast<size_t>(r) * dim_ + c] = tmp[at];
                            if (dbg->field_y) dbg->field_y[static_cast<size_t>(r) * dim_ + c] = tmp[at + 1];
                        }
                }
                fetchT(dbg->force, dForce, 2 * static_cast<size_t>(n_));
                fetchT(dbg->grad_wl, dWl, 2 * static_cast<size_t>(n_));
                fetchT(dbg->grad, g_, 2 * static_cast<size_t>(n_));
                SPHIP_CHECK(hipStreamSynchronize(s_));
            }
        } catch (...) {
            for (void* p : {static_cast<void*>(dNode), static_cast<void*>(dFill),
                            static_cast<void*>(dForce), static_cast<void*>(dWl)})
                if (p) (void)hipFree(p);
            throw;
        }
        for (void* p : {static_cast<void*>(dNode), static_cast<void*>(dFill),
                        static_cast<void*>(dForce), static_cast<void*>(dWl)})
            if (p) SPHIP_CHECK(hipFree(p));
    }

    void run(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows) override;

    bool begin(sp_hip_gp_result* res) override {
        if (cfg_.loop_mode == 0) throw InvalidArg("interleaved stages need loop_mode 1 or 2");
        if (profMask_) throw InvalidArg("interleaved stages are not profiled");
        asyncStart_ = true;
        asyncReady_ = false;
        try {
            run(res, nullptr, 0);
        } catch (...) {
            asyncStart_ = false;
            throw;
        }
        asyncStart_ = false;
        return !asyncReady_;
    }
    void launchAsync() override;
    int pollAsync(bool wait) override;
    int inFlight() const override { return inflight_; }
    void finishAsync(sp_hip_gp_result* res) override {
        if (!asyncReady_) throw InvalidArg("no stage loop is running");
        asyncReady_ = false;
        finishDevice(res, nullptr, 0);
    }

    // The instance's share of a batched slot, after begin() left its device
    // loop ready: the slot kernels' arguments, the Poisson buffers, and the
    // instance's block count in every batched kernel (the counts of its own
    // launches, so that every block does the same work).
    void batchShare(SlotArgs<T>* a, PoissonBatchArgs<T>* p, int counts[BK_KINDS],
                    int* doneOut) {
        if (!asyncReady_) throw InvalidArg("batched slot: the stage loop is not ready");
        if (!poisson_->usesFft())
            throw InvalidArg("batched runs need power-of-two bin grids (FFT Poisson path)");
        const int* skip = reinterpret_cast<const int*>(reinterpret_cast<const char*>(dCtl_) +
                                                       offsetof(Ctl, done));
        const TileGrid& tg = dd_.tiles;
        counts[BK_SORT_COUNT] = static_cast<int>(
            std::min<long>((dd_.nDens + kBlock - 1) / kBlock, kSortBlocks));
        counts[BK_SORT_PLACE] = counts[BK_SORT_COUNT] + 1;   // + the publishing block
        counts[BK_SCATTER] = tg.tilesX * tg.tilesY + (dd_.nDens + tg.chunk - 1) / tg.chunk;
        counts[BK_FINALIZE] = static_cast<int>(
            (std::max<long>(static_cast<long>(nb_), tg.nKeys) + kBlock - 1) / kBlock);
        counts[BK_WA] = dd_.waBlocks;
        counts[BK_NODE] = nodeBlocks(dd_.nCells, n_);
        counts[BK_REDUCE] = 1;
        counts[BK_UPDATE] = counts[BK_NODE];
        if (dd_.nDens == 0) counts[BK_SORT_COUNT] = counts[BK_SORT_PLACE] = 0;
        ReduceWs ws = ws_;
        ws.blocks[0] = counts[BK_FINALIZE];
        ws.blocks[1] = counts[BK_WA];
        ws.blocks[2] = counts[BK_SCATTER];
        ws.blocks[3] = counts[BK_NODE];
        a->d = dd_;
        a->g = geom_;
        a->qscale = static_cast<T>(std::ldexp(1.0, fracBits_));
        a->nodeQ = nodeQ_;
        a->fillQ = fillQ_;
        a->din = DensityIn<T>{geom_, dd_.tiles, nodeQ_, fillQ_, terminal_, base_,
                              std::ldexp(1.0, -fracBits_), binArea_, rhoShift_, ws.region(0),
                              nullptr, nullptr};
        a->rho = poisson_->rho();
        a->exy = poisson_->fieldXY();
        a->ws = ws;
        a->ctl = dCtl_;
        a->rows = dRows_;
        a->bufs = bufTab_;
        a->denom = movableAreaScaled_;
        a->wa = WaTask<T>{dd_, T(0), cfg_.ignore_net_degree, ws.region(1), dCtl_};
        a->doneOut = doneOut;
        *p = poisson_->batchArgs(skip);
    }
    int gridDim() const { return dim_; }
    int sortPeriod() const { return sh_.sort_period; }
    int gridLd() const { return poisson_->ld(); }
    long maxLoopLaunches(int slots) const {
        const int maxIter = cfg_.stage == 1 ? cfg_.filler_only_iters : cfg_.gp_max_iter;
        const int maxBt = cfg_.max_backtrack > 0 ? cfg_.max_backtrack : 10;
        return (static_cast<long>(maxIter) * maxBt + 16) / std::max(1, slots) + 2;
    }
    size_t deviceBytes() const override { return devBytes_; }
    bool batchable() const override { return sizeof(T) == 4 && poisson_->usesFft(); }

private:
    void finishDevice(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows);
    void runDevice(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows, float lambda,
                   float gamma, double alpha0, double effRef, float targetTau, int maxIter,
                   int maxBt, float stepX);
    void launchSlot(bool events, int slot);
    // node coordinates <-> host arrays (the device keeps them interleaved)
    void downloadXY(T* hx, T* hy) const {
        std::vector<T> xy(2 * static_cast<size_t>(total_));
        SPHIP_CHECK(hipMemcpyAsync(xy.data(), dd_.xy, xy.size() * sizeof(T), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
        for (int i = 0; i < total_; ++i) {
            hx[i] = xy[2 * i];
            hy[i] = xy[2 * i + 1];
        }
    }
    void uploadXY(const T* hx, const T* hy) {
        xyStage_.resize(2 * static_cast<size_t>(total_));
        for (int i = 0; i < total_; ++i) {
            xyStage_[2 * i] = hx[i];
            xyStage_[2 * i + 1] = hy[i];
        }
        SPHIP_CHECK(hipMemcpyAsync(dd_.xy, xyStage_.data(), xyStage_.size() * sizeof(T),
                                   hipMemcpyHostToDevice, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }
    std::vector<T> xyStage_;
    // The engine's own node order (nodeOrder), which the caller never sees:
    // ext_ maps the engine's node index to the caller's (empty: the
    // identity); positions and node indices cross the C ABI in the caller's
    // order.
    //  - Fillers carry no pins, so their order is free: they are kept in
    //    Morton order of their start positions (a 1024 x 1024 lattice over
    //    the core), so that neighbouring threads of the node pass and the
    //    update read neighbouring field bins and write neighbouring density
    //    records, and the first spatial sort starts from a nearly sorted order.
    //  - The WA pass gathers a net's pins and the node pass reads their
    //    gradient entries through node indices, so the cells are kept in the
    //    caller's order where its small nets join nearby indices, and
    //    renumbered by breadth-first search over the nets where they do not
    //    (cellsLocal).
    std::vector<int> ext_;
    std::vector<float> ordX_, ordY_, ordW_, ordH_;
    std::vector<uint8_t> ordF_;
    std::vector<int32_t> ordP2N_, ordN2PStart_, ordN2P_;
    int extOf(int i) const { return ext_.empty() ? i : ext_[i]; }
    sp_hip_design nodeOrder(const sp_hip_design& d);
    void build(const sp_hip_design& d);
    void release();
    EvalScalars evaluate(const T* u, const T* up, const T* gp, T* grad, float lambda, float gamma,
                         long long* dbgNode = nullptr, long long* dbgFill = nullptr,
                         T* dbgForce = nullptr, T* dbgWl = nullptr);
    void scatterPos(const T* pos);   // clamp + write into node coordinates
    double deliveredMetrics(double* hpwlOut);
    // HIP event pairs per kernel family, graph slot and kernel of the family
    // (created on demand; evStart_[f][slot * kMaxTimed + k])
    void ensureEvents(int slots) {
        for (int f = 0; f < SP_HIP_K_COUNT; ++f)
            while (static_cast<int>(evStart_[f].size()) < slots * kMaxTimed) {
                hipEvent_t a = nullptr, b = nullptr;
                SPHIP_CHECK(hipEventCreate(&a));
                evStart_[f].push_back(a);
                SPHIP_CHECK(hipEventCreate(&b));
                evStop_[f].push_back(b);
            }
        evUsed_.assign(static_cast<size_t>(SP_HIP_K_COUNT) * slots, 0);
    }
    // The kernels of family f launched between beginK and endK carry the
    // slot's event pairs (launchKernel / hipExtLaunchKernel).
    void beginK(int f, int slot) {
        if (profMask_ & (1 << f)) {
            timer_ = LaunchTimer{};
            for (int k = 0; k < kMaxTimed; ++k) {
                timer_.start[k] = evStart_[f][slot * kMaxTimed + k];
                timer_.stop[k] = evStop_[f][slot * kMaxTimed + k];
            }
            tlsTimer = &timer_;
        }
    }
    void endK(int f, int slot) {
        if (profMask_ & (1 << f)) {
            tlsTimer = nullptr;
            evUsed_[static_cast<size_t>(f) * (evUsed_.size() / SP_HIP_K_COUNT) + slot] = timer_.used;
        }
    }
    LaunchTimer timer_;
    std::vector<int> evUsed_;   // kernels timed per (family, slot) in the last launch of the slot
    // Adds, per profiled family, the summed kernel durations of slots
    // [lo, hi) (one launch of the family per slot that ran it).
    void harvestK(int mask, int lo, int hi) {
        const size_t perF = evUsed_.size() / SP_HIP_K_COUNT;
        for (int f = 0; f < SP_HIP_K_COUNT; ++f) {
            if (!(mask & (1 << f))) continue;
            for (int i = lo; i < hi; ++i) {
                const int used = evUsed_[f * perF + i];
                if (used == 0) continue;   // the family did not run in slot i
                double ms = 0;
                for (int k = 0; k < used; ++k) {
                    float part = 0;
                    SPHIP_CHECK(hipEventSynchronize(evStop_[f][i * kMaxTimed + k]));
                    SPHIP_CHECK(hipEventElapsedTime(&part, evStart_[f][i * kMaxTimed + k],
                                                    evStop_[f][i * kMaxTimed + k]));
                    ms += part;
                }
                kms_[f] += ms;
                klaunch_[f] += 1;
                kbytes_[f] += bytesPer_[f];
                kflops_[f] += flopsPer_[f];
            }
        }
    }
    // A launch outside the loop slots (host-driven evaluations, delivery,
    // position uploads): with family f profiled, its kernels carry event pairs
    // and are harvested at once, so that a profiled run times every launch of
    // the family, as a kernel trace of the same run counts them.
    template <typename F>
    void hostTimed(int f, F&& launch) {
        if (!(profMask_ & (1 << f))) {
            launch();
            return;
        }
        ensureEvents(loopSlots());
        beginK(f, 0);
        launch();
        endK(f, 0);
        harvestK(1 << f, 0, 1);
    }

    sp_hip_gp_config cfg_;
    int device_ = 0;
    hipStream_t s_ = nullptr;   // from the stream pool
    DevDesign<T> dd_{};
    DevArena arena_;            // the design, grids and optimizer state
    sp_hip_shapes sh_{};        // launch shapes for this grid size
    bool fuseWa_ = true;        // column pass and WA blocks in one launch
    void* pinned_ = nullptr;    // pooled pinned block: hScalars_, then hFlag_
    int total_ = 0, n_ = 0, dim_ = 0;
    size_t nb_ = 0;
    double stepX_ = 1, stepY_ = 1, originX_ = 0, originY_ = 0, binArea_ = 1;
    double movableAreaScaled_ = 0;
    double rhoShift_ = 0;   // mean density subtracted before the transforms
    int fracBits_ = 32;
    GridGeom<T> geom_{};
    unsigned long long *nodeQ_ = nullptr, *fillQ_ = nullptr;
    double *terminal_ = nullptr, *base_ = nullptr;
    std::unique_ptr<PoissonPlan<T>> poisson_;
    ReduceWs ws_{};
    double* hScalars_ = nullptr;   // in pinned_
    std::vector<int> optIdxHost_;
    std::vector<T> wHost_, hHost_;
    T *u_ = nullptr, *v_ = nullptr, *uPrev_ = nullptr, *vPrev_ = nullptr, *uNew_ = nullptr,
      *vNew_ = nullptr, *g_ = nullptr, *gPrev_ = nullptr, *gA_ = nullptr, *gB_ = nullptr,
      *probe_ = nullptr;
    const T* curPos_ = nullptr;
    std::vector<T*> snapBufs_;
    // device-resident loop: state, buffer table, metrics rows, snapshot pool, graph
    Ctl* dCtl_ = nullptr;
    BufTable<T> bufTab_{};   // the state buffers of the device loop, by role index
    DevRow* dRows_ = nullptr;
    int rowsCap_ = 0;
    T* snapPool_[kMaxSnaps] = {};
    int* hFlag_ = nullptr;   // in pinned_: copies of ctl->done
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graphExec_ = nullptr;
    int profMask_ = 0;   // bit f: time kernel family f with HIP events
    // the device loop in flight: launches enqueued, each followed by a copy
    // of ctl->done into hFlag_[slot] and an event (a ring of kAsyncDepth)
    bool asyncStart_ = false, asyncReady_ = false, loopGraph_ = false;
    int loopSlots_ = 1, inflight_ = 0, ringHead_ = 0;
    long loopLaunches_ = 0, maxLaunches_ = 0;
    std::chrono::steady_clock::time_point loopT0_{};
    hipEvent_t loopEv_[kAsyncDepth] = {};
    size_t devBytes_ = 0;
    std::vector<hipEvent_t> evStart_[SP_HIP_K_COUNT], evStop_[SP_HIP_K_COUNT];
    double kms_[SP_HIP_K_COUNT] = {}, kbytes_[SP_HIP_K_COUNT] = {}, kflops_[SP_HIP_K_COUNT] = {};
    int64_t klaunch_[SP_HIP_K_COUNT] = {};
    double bytesPer_[SP_HIP_K_COUNT] = {}, flopsPer_[SP_HIP_K_COUNT] = {};
};

template <typename T>
void Engine<T>::release() {
    if (s_) (void)hipStreamSynchronize(s_);
    poisson_.reset();
    arena_.release();
    for (T* p : snapBufs_) (void)hipFree(p);
    snapBufs_.clear();
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graphExec_ = nullptr;
    graph_ = nullptr;
    if (dRows_) (void)hipFree(dRows_);
    dRows_ = nullptr;
    for (hipEvent_t& ev : loopEv_) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    givePinned(pinned_);
    pinned_ = nullptr;
    hScalars_ = nullptr;
    hFlag_ = nullptr;
    for (int f = 0; f < SP_HIP_K_COUNT; ++f) {
        for (hipEvent_t ev : evStart_[f]) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : evStop_[f]) (void)hipEventDestroy(ev);
        evStart_[f].clear();
        evStop_[f].clear();
    }
    giveStream(device_, s_);
    s_ = nullptr;
}

namespace {

// Nets of 2 to 4 pins whose movable pins span more than this many node
// indices gather from separate cache lines in the WA pass.
constexpr int kLocalSpan = 256;

// True when at most a tenth of the design's 2-4 pin nets span more than
// kLocalSpan movable node indices (nets built from a hierarchy, or numbered
// by connectivity, as the benchmark generator's are); false for an order
// that scatters nets over the index range.
bool cellsLocal(const sp_hip_design& d) {
    const int nm = d.num_movable;
    long small = 0, far = 0;
    for (int e = 0; e < d.num_nets; ++e) {
        const int p0 = d.net2pin_start[e], p1 = d.net2pin_start[e + 1];
        if (p1 - p0 < 2 || p1 - p0 > 4) continue;
        int lo = nm, hi = -1;
        for (int p = p0; p < p1; ++p) {
            const int i = d.pin2node[d.flat_net2pin[p]];
            if (i < nm) {
                lo = std::min(lo, i);
                hi = std::max(hi, i);
            }
        }
        if (hi < lo) continue;
        ++small;
        if (hi - lo > kLocalSpan) ++far;
    }
    return 10 * far <= small;
}

// Movable nodes in breadth-first order over the nets (nets above 16 pins
// carry no locality worth following and are skipped), from the lowest
// unvisited index each time the queue empties: the nodes of a net are
// numbered together when its first node is reached.  order[new] = old.
std::vector<int> bfsCells(const sp_hip_design& d) {
    constexpr int kMaxDeg = 16;
    const int nm = d.num_movable;
    std::vector<int> order;
    order.reserve(nm);
    std::vector<char> seen(nm, 0);
    for (int s = 0; s < nm; ++s) {
        if (seen[s]) continue;
        seen[s] = 1;
        size_t head = order.size();
        order.push_back(s);
        while (head < order.size()) {
            const int u = order[head++];
            for (int q = d.node2pin_start[u]; q < d.node2pin_start[u + 1]; ++q) {
                const int e = d.pin2net[d.flat_node2pin[q]];
                const int p0 = d.net2pin_start[e], p1 = d.net2pin_start[e + 1];
                if (p1 - p0 > kMaxDeg) continue;
                for (int p = p0; p < p1; ++p) {
                    const int v = d.pin2node[d.flat_net2pin[p]];
                    if (v < nm && !seen[v]) {
                        seen[v] = 1;
                        order.push_back(v);
                    }
                }
            }
        }
    }
    return order;
}

}  // namespace

template <typename T>
sp_hip_design Engine<T>::nodeOrder(const sp_hip_design& d) {
    const int nn = d.num_nodes, nm = d.num_movable, total = d.num_nodes + d.num_fillers;
    const bool renumber = nm > 1 && !cellsLocal(d);
    if (d.num_fillers < 2 && !renumber) return d;
    ext_.resize(total);
    std::iota(ext_.begin(), ext_.end(), 0);
    if (renumber) {
        const std::vector<int> order = bfsCells(d);
        std::copy(order.begin(), order.end(), ext_.begin());
    }
    if (d.num_fillers >= 2) {
        const double lx = d.core_lx, ly = d.core_ly;
        const double w = std::max(1e-30, static_cast<double>(d.core_hx) - lx);
        const double h = std::max(1e-30, static_cast<double>(d.core_hy) - ly);
        auto lattice = [](double v) {
            return static_cast<uint32_t>(std::clamp(static_cast<int>(v * 1024.0), 0, 1023));
        };
        auto morton = [](uint32_t ix, uint32_t iy) {
            uint64_t k = 0;
            for (int b = 0; b < 10; ++b)
                k |= static_cast<uint64_t>(((ix >> b) & 1u) | (((iy >> b) & 1u) << 1)) << (2 * b);
            return k;
        };
        std::vector<uint64_t> key(d.num_fillers);
        for (int f = 0; f < d.num_fillers; ++f)
            key[f] = morton(lattice((d.node_x[nn + f] - lx) / w),
                            lattice((d.node_y[nn + f] - ly) / h));
        std::stable_sort(ext_.begin() + nn, ext_.end(),
                         [&](int a, int b) { return key[a - nn] < key[b - nn]; });
    }
    auto gather = [&](auto& dst, const auto* src) {
        dst.resize(total);
        for (int i = 0; i < total; ++i) dst[i] = src[ext_[i]];
        return dst.data();
    };
    sp_hip_design p = d;
    p.node_x = gather(ordX_, d.node_x);
    p.node_y = gather(ordY_, d.node_y);
    p.node_w = gather(ordW_, d.node_w);
    p.node_h = gather(ordH_, d.node_h);
    p.node_flags = gather(ordF_, d.node_flags);
    if (renumber) {
        // pins keep their ids; each node keeps its pins in the caller's order
        std::vector<int> inv(total);
        for (int i = 0; i < total; ++i) inv[ext_[i]] = i;
        ordP2N_.resize(d.num_pins);
        for (int q = 0; q < d.num_pins; ++q) ordP2N_[q] = inv[d.pin2node[q]];
        ordN2PStart_.assign(total + 1, 0);
        ordN2P_.resize(d.node2pin_start[total]);
        for (int i = 0; i < total; ++i) {
            const int o = ext_[i];
            const int c = d.node2pin_start[o + 1] - d.node2pin_start[o];
            ordN2PStart_[i + 1] = ordN2PStart_[i] + c;
            std::copy(d.flat_node2pin + d.node2pin_start[o], d.flat_node2pin + d.node2pin_start[o + 1],
                      ordN2P_.begin() + ordN2PStart_[i]);
        }
        p.pin2node = ordP2N_.data();
        p.node2pin_start = ordN2PStart_.data();
        p.flat_node2pin = ordN2P_.data();
    }
    return p;
}

template <typename T>
void Engine<T>::build(const sp_hip_design& d) {
    const int total = d.num_nodes + d.num_fillers;
    total_ = total;
    const int stage = cfg_.stage;
    auto isMacro = [&](int i) { return (d.node_flags[i] & F_MACRO) != 0; };
    auto isFixed = [&](int i) { return (d.node_flags[i] & F_FIXED) != 0; };
    auto isNI = [&](int i) { return (d.node_flags[i] & F_NI) != 0; };

    // buildOptSet, eplace.cpp:82-94.
    if (stage != 1)
        for (int i = 0; i < d.num_movable; ++i) {
            if (isFixed(i) || isNI(i)) continue;
            if (stage == 2 && isMacro(i)) continue;
            optIdxHost_.push_back(i);
        }
    dd_.nCells = static_cast<int>(optIdxHost_.size());
    dd_.cellBlocks = nodePartBlocks(dd_.nCells);
    for (int i = d.num_nodes; i < total; ++i) optIdxHost_.push_back(i);
    n_ = static_cast<int>(optIdxHost_.size());

    // BinGrid::initialize, bin_grid.cpp:172-281.
    const float target = cfg_.target_density;
    double cellArea = 0, macroArea = 0;
    movableAreas(d, &cellArea, &macroArea);
    movableAreaScaled_ = cellArea + macroArea * static_cast<double>(target);
    dim_ = binDim(d, cfg_);
    sh_ = shapesFor(dim_);
    if (dim_ < 2) throw InvalidArg("bin grid must be at least 2x2");
    const float coreW = d.core_hx - d.core_lx, coreH = d.core_hy - d.core_ly;
    if (sizeof(T) == 4) {
        stepX_ = coreW / static_cast<float>(dim_);
        stepY_ = coreH / static_cast<float>(dim_);
    } else {
        stepX_ = static_cast<double>(coreW) / dim_;
        stepY_ = static_cast<double>(coreH) / dim_;
    }
    originX_ = d.core_lx;
    originY_ = d.core_ly;
    binArea_ = stepX_ * stepY_;
    nb_ = static_cast<size_t>(dim_) * dim_;

    // static terminal density (bin_grid.cpp:228-245) and base density (:252-281),
    // in the working precision
    const T sx = T(stepX_), sy = T(stepY_), ox = T(originX_), oy = T(originY_);
    std::vector<T> term(nb_, T(0)), base(nb_, T(0));
    for (int i = 0; i < d.num_nodes; ++i) {
        if (!isFixed(i) || isNI(i)) continue;
        const T lx = T(d.node_x[i]) - T(0.5) * T(d.node_w[i]);
        const T hx = T(d.node_x[i]) + T(0.5) * T(d.node_w[i]);
        const T ly = T(d.node_y[i]) - T(0.5) * T(d.node_h[i]);
        const T hy = T(d.node_y[i]) + T(0.5) * T(d.node_h[i]);
        auto cl = [&](T v) { return std::clamp(static_cast<int>(v), 0, dim_ - 1); };
        const int bx0 = cl((lx - ox) / sx), bx1 = cl((hx - ox) / sx);
        const int by0 = cl((ly - oy) / sy), by1 = cl((hy - oy) / sy);
        for (int bx = bx0; bx <= bx1; ++bx) {
            const T blx = ox + T(bx) * sx;
            for (int by = by0; by <= by1; ++by) {
                const T bly = oy + T(by) * sy;
                term[static_cast<size_t>(bx) * dim_ + by] +=
                    T(target) * overlapT<T>(blx, bly, blx + sx, bly + sy, lx, ly, hx, hy);
            }
        }
    }
    std::vector<int> order(d.num_rows);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return d.row_ly[a] < d.row_ly[b]; });
    for (int by = 0; by < dim_; ++by) {
        const T bly = oy + T(by) * sy, bhy = bly + sy;
        size_t lo = 0, hi = order.size();
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            const int r = order[mid];
            if (T(d.row_ly[r]) + T(d.row_height[r]) <= bly) lo = mid + 1;
            else hi = mid;
        }
        for (int bx = 0; bx < dim_; ++bx) {
            const T blx = ox + T(bx) * sx;
            T covered = 0;
            for (size_t t = lo; t < order.size(); ++t) {
                const int r = order[t];
                if (T(d.row_ly[r]) >= bhy) break;
                covered += overlapT<T>(blx, bly, blx + sx, bhy, T(d.row_lx[r]), T(d.row_ly[r]),
                                       T(d.row_hx[r]), T(d.row_ly[r]) + T(d.row_height[r]));
            }
            const T freeA = std::max(T(0), (blx + sx - blx) * (bhy - bly) - covered);
            base[static_cast<size_t>(bx) * dim_ + by] = T(target) * freeA;
        }
    }
    arena_.upload(&terminal_, term);
    arena_.upload(&base_, base);

    // fixed-point scale: the largest possible bin total (all density charge in
    // one bin) stays below 2^61
    std::vector<int> dens;
    double chargeBound = 1.0;
    for (int i = 0; i < total; ++i) {
        if (isFixed(i) || isNI(i)) continue;
        dens.push_back(i);
        chargeBound += static_cast<double>(d.node_w[i]) * d.node_h[i] *
                       std::max(1.0, static_cast<double>(target));
    }
    fracBits_ = std::min(40, 61 - static_cast<int>(std::ceil(std::log2(chargeBound))));
    {
        double charge = 0;
        for (int i : dens) {
            const double a = static_cast<double>(d.node_w[i]) * d.node_h[i];
            charge += ((d.node_flags[i] & F_MACRO) && !(d.node_flags[i] & F_FILLER))
                          ? a * static_cast<double>(target) : a;
        }
        for (size_t b = 0; b < nb_; ++b) charge += static_cast<double>(term[b]) + base[b];
        rhoShift_ = charge / (binArea_ * static_cast<double>(nb_));
    }
    if (fracBits_ < 8) throw InvalidArg("design too large for the fixed-point density grid");
    geom_.dim = dim_;
    geom_.stepX = T(stepX_);
    geom_.stepY = T(stepY_);
    geom_.originX = T(originX_);
    geom_.originY = T(originY_);
    geom_.target = T(target);

    poisson_ = std::make_unique<PoissonPlan<T>>(dim_, dim_, stepX_, stepY_, s_, true, &arena_);
    geom_.ld = poisson_->ld();
    // a column pass of its own where the shapes ask for one and the plan takes it
    fuseWa_ = !(sh_.col_pairs > 0 && poisson_->setColumnPairs(sh_.col_pairs));

    // device design
    dd_.total = total;
    dd_.numMov = d.num_movable;
    dd_.numNodes = d.num_nodes;
    dd_.nOpt = n_;
    dd_.nDens = static_cast<int>(dens.size());
    dd_.numNets = d.num_nets;
    dd_.numPins = d.num_pins;
    dd_.coreLx = T(d.core_lx);
    dd_.coreLy = T(d.core_ly);
    dd_.coreHx = T(d.core_hx);
    dd_.coreHy = T(d.core_hy);
    std::vector<float> xs(d.node_x, d.node_x + total), ys(d.node_y, d.node_y + total);
    {
        std::vector<float> xy(2 * static_cast<size_t>(total));
        for (int i = 0; i < total; ++i) {
            xy[2 * i] = xs[i];
            xy[2 * i + 1] = ys[i];
        }
        arena_.upload(&dd_.xy, xy);
    }
    std::vector<float> ws(d.node_w, d.node_w + total), hs(d.node_h, d.node_h + total);
    wHost_.assign(ws.begin(), ws.end());
    hHost_.assign(hs.begin(), hs.end());
    arena_.upload(&dd_.w, ws);
    arena_.upload(&dd_.h, hs);
    std::vector<uint8_t> fl(d.node_flags, d.node_flags + total);
    arena_.upload(&dd_.flags, fl);
    arena_.upload(&dd_.optIdx, optIdxHost_);
    {
        // optIdx = k below a split and k + shift from it on: the kernels compute it
        int split = 0;
        while (split < n_ && optIdxHost_[split] == split) ++split;
        const int shift = split < n_ ? optIdxHost_[split] - split : 0;
        bool affine = true;
        for (int k = split; k < n_ && affine; ++k) affine = optIdxHost_[k] == k + shift;
        dd_.optAffine = affine ? 1 : 0;
        dd_.optSplit = split;
        dd_.optShift = shift;
    }
    {
        // density records in node order to start with; the first sort orders them
        std::vector<int> recId(dens.size());
        std::vector<float> recWH(2 * dens.size());
        for (size_t t = 0; t < dens.size(); ++t) {
            const int i = dens[t];
            if (i > kRecNode) throw InvalidArg("too many nodes");
            recId[t] = i | ((d.node_flags[i] & F_FILLER) ? kRecFiller : 0) |
                       ((d.node_flags[i] & F_MACRO) ? kRecMacro : 0);
            recWH[2 * t] = d.node_w[i];
            recWH[2 * t + 1] = d.node_h[i];
        }
        arena_.upload(&dd_.recIdA, recId);
        arena_.upload(&dd_.recIdB, recId);
        arena_.upload(&dd_.recWHA, recWH);
        arena_.upload(&dd_.recWHB, recWH);
    }
    {
        TileGrid& tg = dd_.tiles;
        tg.tilesX = tg.tilesY = (dim_ + kTileBins - 1) / kTileBins;
        int side = 1;   // Morton keys of tiles below side x side, side a power of two
        while (side < tg.tilesX) side <<= 1;
        tg.nKeys = side * side;
        if (tg.nKeys > kTileKeys) throw InvalidArg("bin grid wider than 1024 bins");
        tg.chunk = sh_.scatter_chunk;
        arena_.zeros(&tg.count, tg.nKeys);
        arena_.zeros(&tg.fill, tg.nKeys);
        arena_.zeros(&tg.start, tg.nKeys);
        arena_.zeros(&tg.size, tg.nKeys);
        arena_.zeros(&tg.extraStart, tg.nKeys + 1);
        const size_t stageLen = static_cast<size_t>(tg.tilesX) * tg.tilesY * 2 * kTileWin * kTileWin;
        arena_.zeros(&tg.stage, stageLen);
    }
    // Nets are renumbered by WA size class (degree 2, 3-4, 5-8, ..., then
    // larger), so each class is a contiguous range of nets and of pin slots,
    // and inside a class by their lowest node index, so that neighbouring
    // nets gather coordinates of neighbouring nodes (netlists are local in
    // node order).  Pin slots are laid out in the new net order.
    auto degOf = [&](int k) { return d.net2pin_start[k + 1] - d.net2pin_start[k]; };
    auto classOf = [&](int deg) {
        int c = 0;
        while (c < kWaClasses && deg > (2 << c)) ++c;
        return deg < 2 ? -1 : c;
    };
    std::vector<int> netMinNode(d.num_nets, 0x7FFFFFFF);
    for (int k = 0; k < d.num_nets; ++k)
        for (int q = d.net2pin_start[k]; q < d.net2pin_start[k + 1]; ++q)
            netMinNode[k] = std::min(netMinNode[k], d.pin2node[d.flat_net2pin[q]]);
    std::vector<int> netOrder(d.num_nets);
    std::iota(netOrder.begin(), netOrder.end(), 0);
    std::stable_sort(netOrder.begin(), netOrder.end(), [&](int a, int b) {
        const int ca = classOf(degOf(a)), cb = classOf(degOf(b));
        if (ca != cb) return ca < cb;
        if (ca == kWaClasses) return degOf(a) < degOf(b);
        return netMinNode[a] < netMinNode[b];
    });
    std::vector<int> netStart(d.num_nets + 1, 0);
    std::vector<int> slotNode(d.num_pins), pinSlot(d.num_pins, -1);
    std::vector<float> slotOffX(d.num_pins), slotOffY(d.num_pins);
    {
        int t = 0;
        for (int k = 0; k < d.num_nets; ++k) {
            const int old = netOrder[k];
            netStart[k] = t;
            for (int q = d.net2pin_start[old]; q < d.net2pin_start[old + 1]; ++q, ++t) {
                const int p = d.flat_net2pin[q];
                slotNode[t] = d.pin2node[p];
                slotOffX[t] = d.pin_offset_x[p];
                slotOffY[t] = d.pin_offset_y[p];
                pinSlot[p] = t;
            }
        }
        netStart[d.num_nets] = t;
    }
    // size classes: first net of degree > 2^c (class c holds (2^c, 2^(c+1)],
    // class 0 holds degree 2), wavefronts per class
    {
        // class boundaries in the new order (classes are contiguous, ascending)
        std::vector<int> cls(d.num_nets);
        for (int k = 0; k < d.num_nets; ++k) cls[k] = classOf(netStart[k + 1] - netStart[k]);
        for (int c = 0; c <= kWaClasses; ++c) {
            int k = 0;
            while (k < d.num_nets && cls[k] < c) ++k;
            dd_.waNetLo[c] = k;
        }
        dd_.waNetLo[kWaClasses + 1] = d.num_nets;
        dd_.waWaveBase[0] = 0;
        dd_.waLaneBase[0] = 0;
        for (int c = 0; c < kWaClasses; ++c) {
            const int nets = dd_.waNetLo[c + 1] - dd_.waNetLo[c];
            const int tasks = (nets + waNetsPerTask(c) - 1) / waNetsPerTask(c);
            dd_.waWaveBase[c + 1] = dd_.waWaveBase[c] + tasks;
            dd_.waLaneBase[c + 1] = dd_.waLaneBase[c] + static_cast<long>(tasks) * waEntriesPerTask(c);
        }
        dd_.waWaveBase[kWaClasses + 1] =
            dd_.waWaveBase[kWaClasses] + (dd_.waNetLo[kWaClasses + 1] - dd_.waNetLo[kWaClasses]);
        // WA blocks: no more than the GPU holds at once, in the WA pass's own
        // launch and, where the loop fuses it, beside the column blocks (a
        // block that waits for a free slot starts a second round and sets the
        // launch's length); every WA pass of the instance takes this count,
        // so its partials sum in one order
        int resident = waResidentBlocks<T>();
        if (fuseWa_) {
            const int fused = poisson_->fusedWaResidentBlocks();
            if (fused > 0) resident = std::min(resident, fused);
        }
        dd_.waBlocks = waBlocksFor(dd_.waWaveBase[kWaClasses + 1], resident);
        // tasks by XCD: the XCD that runs the node-pass block of the lowest
        // node of the task's first net (block b of a part runs logical block
        // xcdBlock(b, part) on XCD b % kXcds), each XCD's tasks longest first
        // (class order reversed); nodes outside the optimised set by index
        const int tasks = dd_.waWaveBase[kWaClasses + 1];
        const int cellShare = std::max(1, dd_.cellBlocks / kXcds);
        const int fillShare = std::max(1, nodePartBlocks(n_ - dd_.nCells) / kXcds);
        std::vector<int> kOfNode(total, -1);
        for (int k = 0; k < n_; ++k) kOfNode[optIdxHost_[k]] = k;
        auto xcdOfNode = [&](int i) {
            const int k = kOfNode[i];
            if (k < 0) return std::min(kXcds - 1, static_cast<int>(static_cast<long>(i) * kXcds / total));
            if (k < dd_.nCells) return std::min(kXcds - 1, (k / kBlock) / cellShare);
            return std::min(kXcds - 1, ((k - dd_.nCells) / kBlock) / fillShare);
        };
        std::vector<std::vector<int>> perXcd(kXcds);
        for (int gw = tasks - 1; gw >= 0; --gw) {
            int c = 0;
            while (c <= kWaClasses && gw >= dd_.waWaveBase[c + 1]) ++c;
            const int first = dd_.waNetLo[c] + (c < kWaClasses ? (gw - dd_.waWaveBase[c]) * waNetsPerTask(c)
                                                               : gw - dd_.waWaveBase[c]);
            const int node = netMinNode[netOrder[std::min(first, d.num_nets - 1)]];
            perXcd[node >= 0 && node < total ? xcdOfNode(node) : 0].push_back(gw);
        }
        std::vector<int> order;
        order.reserve(tasks);
        dd_.waXcdStart[0] = 0;
        for (int x = 0; x < kXcds; ++x) {
            order.insert(order.end(), perXcd[x].begin(), perXcd[x].end());
            dd_.waXcdStart[x + 1] = static_cast<int>(order.size());
        }
        arena_.upload(&dd_.waTaskOrder, order);
    }
    // node-ordered pin entries; slot -> entry for the gradient writes.  Entries
    // of pins that no net pass writes (no net, degree < 2 or above
    // ignore_net_degree) stay zero.
    std::vector<int> nodeStart(d.node2pin_start, d.node2pin_start + total + 1);
    std::vector<int> slotNp(d.num_pins, 0);
    for (int q = 0; q < d.num_pins; ++q) {
        const int p = d.flat_node2pin[q];
        if (pinSlot[p] >= 0) slotNp[pinSlot[p]] = q;
    }
    arena_.upload(&dd_.nodeStart, nodeStart);
    // classes below kWaClasses in lane order, from entry waLaneBase[c]:
    //  - c < kWaLaneClasses: net g is lane g % 64 of task g / 64, its pin j
    //    entry (task P + j) 64 + lane (P = 2^(c+1) pins per lane);
    //  - otherwise net g (W = 2^(c+1) lanes per net, 64 / W nets per task)
    //    takes lanes (g % (64 / W)) W + j, j < degree, of task g / (64 / W)
    {
        const long lanes = dd_.waLaneBase[kWaClasses];
        std::vector<int> laneNode(lanes, -1), laneNp(lanes, 0);
        std::vector<float> laneOffX(lanes, 0.f), laneOffY(lanes, 0.f);
        for (int c = 0; c < kWaClasses; ++c) {
            const int W = 2 << c, perWave = 64 / W;
            for (int k = dd_.waNetLo[c]; k < dd_.waNetLo[c + 1]; ++k) {
                const int g = k - dd_.waNetLo[c];
                for (int t = netStart[k]; t < netStart[k + 1]; ++t) {
                    const int j = t - netStart[k];
                    const long e = c < kWaLaneClasses
                        ? dd_.waLaneBase[c] + (static_cast<long>(g / 64) * W + j) * 64 + g % 64
                        : dd_.waLaneBase[c] + 64L * (g / perWave) + (g % perWave) * W + j;
                    laneNode[e] = slotNode[t];
                    laneOffX[e] = slotOffX[t];
                    laneOffY[e] = slotOffY[t];
                    laneNp[e] = slotNp[t];
                }
            }
        }
        arena_.upload(&dd_.laneNode, laneNode);
        arena_.upload(&dd_.laneOffX, laneOffX);
        arena_.upload(&dd_.laneOffY, laneOffY);
        arena_.upload(&dd_.laneNp, laneNp);
    }
    // the tail class keeps per-net slot ranges, rebased to its first slot
    {
        const int k0 = dd_.waNetLo[kWaClasses], t0 = netStart[k0];
        std::vector<int> tailStart(netStart.begin() + k0, netStart.end());
        for (int& t : tailStart) t -= t0;
        arena_.upload(&dd_.tailStart, tailStart);
        auto tail = [&](auto& v) { return std::vector<typename std::decay_t<decltype(v)>::value_type>(
                                       v.begin() + t0, v.end()); };
        arena_.upload(&dd_.slotNode, tail(slotNode));
        arena_.upload(&dd_.slotOffX, tail(slotOffX));
        arena_.upload(&dd_.slotOffY, tail(slotOffY));
        arena_.upload(&dd_.slotNp, tail(slotNp));
    }
    arena_.zeros(&dd_.pinGrad, 2 * static_cast<size_t>(std::max(1, d.num_pins)));

    arena_.zeros(&nodeQ_, nb_);
    arena_.zeros(&fillQ_, nb_);

    // reduction workspace: 4 regions of maxBlocks x 8 doubles (one block
    // partial per region, see ReduceWs)
    const long maxBlocks = std::max<long>(
        {(static_cast<long>(nb_) + kBlock - 1) / kBlock, (dd_.waWaveBase[kWaClasses + 1] + 3L) / 4,
         static_cast<long>(nodeBlocks(dd_.nCells, n_)),
         static_cast<long>(dd_.tiles.tilesX) * dd_.tiles.tilesY +
             (dd_.nDens + dd_.tiles.chunk - 1) / dd_.tiles.chunk,
         1L});
    ws_.maxBlocks = static_cast<int>(maxBlocks) + 1;
    arena_.zeros(&ws_.partials, 4L * ws_.maxBlocks * 8);
    arena_.zeros(&ws_.scalars, kNumScalars);
    static_assert(kNumScalars * sizeof(double) + kAsyncDepth * sizeof(int) <= pools::kPinnedBlock,
                  "the pinned block holds the scalars and the done flags");
    pinned_ = takePinned();
    std::memset(pinned_, 0, pools::kPinnedBlock);
    hScalars_ = static_cast<double*>(pinned_);
    hFlag_ = reinterpret_cast<int*>(static_cast<char*>(pinned_) + kNumScalars * sizeof(double));

    // optimizer state
    const size_t n2 = 2 * static_cast<size_t>(std::max(n_, 1));
    arena_.zeros(&dCtl_, 1);
    if (cfg_.loop_mode != 0)
        for (int i = 0; i < kMaxSnaps; ++i) arena_.zeros(&snapPool_[i], n2);
    for (hipEvent_t& ev : loopEv_) SPHIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    for (T** b : {&u_, &v_, &uPrev_, &vPrev_, &uNew_, &vNew_, &g_, &gPrev_, &gA_, &gB_, &probe_})
        arena_.zeros(b, n2);
    arena_.place(s_);

    // algorithmic bytes / flops per launch (the roofline accounting)
    const double P = d.num_pins, ts = sizeof(T), nd = dd_.nDens;
    const double stageBytes = static_cast<double>(dd_.tiles.tilesX) * dd_.tiles.tilesY * 2 *
                              kTileWin * kTileWin * 8;
    // sort: pass 1 reads each record's node index and gathers its (x, y); pass
    // 2 reads the record (index, w/h) and (x, y) again and writes the record
    // in the new order
    bytesPer_[SP_HIP_K_SORT] = nd * ((4 + 2 * ts) + (4 + 2 * ts + 2 * ts + 4 + 2 * ts));
    // scatter: reads the ordered records and their (x, y), writes the records
    // back and every tile window
    bytesPer_[SP_HIP_K_SCATTER] = nd * (4 + 2 * ts + 2 * ts + 4 + 2 * ts) + stageBytes;
    // finalize: atomic grids read and cleared, windows read, static terrain, rho
    bytesPer_[SP_HIP_K_FINALIZE] = nb_ * (8.0 * 2 + 8 * 2 + 8 * 2 + ts) + stageBytes;
    bytesPer_[SP_HIP_K_POISSON] = poisson_->bytes(false);
    if (poisson_->takesDensity()) {
        // the first pass finalizes the grids: the finalize bytes join the
        // solve's, less rho's write and read
        bytesPer_[SP_HIP_K_POISSON] += bytesPer_[SP_HIP_K_FINALIZE] - 2.0 * nb_ * ts;
        bytesPer_[SP_HIP_K_FINALIZE] = 0;
    }
    flopsPer_[SP_HIP_K_POISSON] = poisson_->flops(false);
    // WA net pass: per lane of the class tasks its node id, offsets and entry
    // id; per pin the node's x/y and the gradient write; tail slot ranges
    bytesPer_[SP_HIP_K_WA_NETS] = static_cast<double>(dd_.waLaneBase[kWaClasses]) * (4 + 2 * ts + 4) +
                                  P * (2 * ts + 2 * ts) +
                                  (dd_.waNetLo[kWaClasses + 1] - dd_.waNetLo[kWaClasses]) * 4.0;
    // node pass: per node its index, x/y, w/h, flags, pin range (two ints),
    // u/up/gp reads and the grad write, four field bins (x and y); per pin
    // entry its gradient
    bytesPer_[SP_HIP_K_NODE_GRAD] =
        n_ * (4 + 2 * ts + 2 * ts + 1 + 4 + 4 + 4 * 2 * ts) + P * 2 * ts + n_ * 4.0 * 2 * ts;
    // update: w/h, A/G/B reads, v/u writes, the node's (x, y) write
    bytesPer_[SP_HIP_K_UPDATE] = n_ * (2 * ts + 3 * 2 * ts + 2 * 2 * ts + 2 * ts);
    // reduce: the block partials of the evaluation (1, 2 and 7 doubles)
    bytesPer_[SP_HIP_K_REDUCE] =
        8.0 * ((nb_ + kBlock - 1) / kBlock +
               2.0 * dd_.waBlocks +
               7.0 * nodeBlocks(dd_.nCells, n_));
}

template <typename T>
EvalScalars Engine<T>::evaluate(const T* u, const T* up, const T* gp, T* grad, float lambda,
                                float gamma, long long* dbgNode, long long* dbgFill, T* dbgForce,
                                T* dbgWl) {
    hostTimed(SP_HIP_K_SORT, [&] { launchTileSort<T>(dd_, geom_, nullptr, s_); });
    hostTimed(SP_HIP_K_SCATTER, [&] {
        launchScatter<T>(dd_, geom_, static_cast<T>(std::ldexp(1.0, fracBits_)), nodeQ_, fillQ_,
                         true, ws_, nullptr, s_);
    });
    hostTimed(SP_HIP_K_FINALIZE, [&] {
        launchFinalize<T>(geom_, dd_.tiles, nodeQ_, fillQ_, terminal_, base_,
                          std::ldexp(1.0, -fracBits_), binArea_, rhoShift_, poisson_->rho(), dbgNode,
                          dbgFill, ws_, nullptr, s_);
    });
    hostTimed(SP_HIP_K_POISSON, [&] { poisson_->solve(false); });
    hostTimed(SP_HIP_K_WA_NETS, [&] {
        launchWaNets<T>(dd_, gamma, cfg_.ignore_net_degree, ws_, nullptr, s_);
    });
    hostTimed(SP_HIP_K_NODE_GRAD, [&] {
        launchNodeGrad<T>(dd_, geom_, poisson_->fieldXY(), gamma, lambda, u, up, gp, grad,
                          dbgForce, dbgWl, ws_, nullptr, BufTable<T>{}, s_);
    });
    hostTimed(SP_HIP_K_REDUCE, [&] {
        launchReduce<T>(ws_, movableAreaScaled_, true, nullptr, nullptr, s_);
    });
    SPHIP_CHECK(hipMemcpyAsync(hScalars_, ws_.scalars, kNumScalars * sizeof(double),
                               hipMemcpyDeviceToHost, s_));
    SPHIP_CHECK(hipStreamSynchronize(s_));
    EvalScalars sc;
    std::memcpy(&sc, hScalars_, sizeof(sc));
    return sc;
}

template <typename T>
void Engine<T>::scatterPos(const T* pos) {
    hostTimed(SP_HIP_K_UPDATE, [&] {
        launchUpdate<T>(dd_, pos, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr, nullptr,
                        BufTable<T>{}, s_);
    });
    curPos_ = pos;
}

// Delivery (eplace.cpp:478-494): positions already scattered; tau and HPWL.
template <typename T>
double Engine<T>::deliveredMetrics(double* hpwlOut) {
    hostTimed(SP_HIP_K_SORT, [&] { launchTileSort<T>(dd_, geom_, nullptr, s_); });
    hostTimed(SP_HIP_K_SCATTER, [&] {
        launchScatter<T>(dd_, geom_, static_cast<T>(std::ldexp(1.0, fracBits_)), nodeQ_, fillQ_,
                         true, ws_, nullptr, s_);
    });
    hostTimed(SP_HIP_K_FINALIZE, [&] {
        launchFinalize<T>(geom_, dd_.tiles, nodeQ_, fillQ_, terminal_, base_,
                          std::ldexp(1.0, -fracBits_), binArea_, rhoShift_, poisson_->rho(), nullptr,
                          nullptr, ws_, nullptr, s_);
    });
    hostTimed(SP_HIP_K_WA_NETS, [&] {
        launchWaNets<T>(dd_, 1.0f, cfg_.ignore_net_degree, ws_, nullptr, s_);
    });
    hostTimed(SP_HIP_K_REDUCE, [&] {
        launchReduce<T>(ws_, movableAreaScaled_, false, nullptr, nullptr, s_);
    });
    SPHIP_CHECK(hipMemcpyAsync(hScalars_, ws_.scalars, kNumScalars * sizeof(double),
                               hipMemcpyDeviceToHost, s_));
    SPHIP_CHECK(hipStreamSynchronize(s_));
    *hpwlOut = hScalars_[2];
    return hScalars_[0];
}

template <typename T>
void Engine<T>::run(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows) {
    *res = sp_hip_gp_result{};
    res->rollback_iter = -1;
    const int stage = cfg_.stage;
    const int n = n_;
    if (n == 0) return;
    const size_t bytes2n = 2 * static_cast<size_t>(n) * sizeof(T);
    const float stepX = static_cast<float>(stepX_);

    // gather -> clamp -> scatter of the initial positions (eplace.cpp:238-245)
    {
        std::vector<T> hx(total_), hy(total_);
        downloadXY(hx.data(), hy.data());
        std::vector<T> pos(2 * static_cast<size_t>(n));
        for (int k = 0; k < n; ++k) {
            pos[k] = hx[optIdxHost_[k]];
            pos[n + k] = hy[optIdxHost_[k]];
        }
        SPHIP_CHECK(hipMemcpyAsync(probe_, pos.data(), bytes2n, hipMemcpyHostToDevice, s_));
        // clamp into u (v' = clamp(A) with vOut = u)
        launchUpdate<T>(dd_, probe_, nullptr, nullptr, 0.0, 0.0, u_, nullptr, nullptr,
                        BufTable<T>{}, s_);
        curPos_ = u_;
    }

    auto checkFinite = [&](const EvalScalars& sc) {
        if (!std::isfinite(sc.tau) || !std::isfinite(sc.hpwl) || !std::isfinite(sc.wl) ||
            !std::isfinite(sc.sum_wl) || !std::isfinite(sc.sum_den) || sc.nonfinite > 0)
            throw NonFinite("GP/evaluate: non-finite objective, overflow, force norm or gradient");
    };

    // lambda initialisation probe, eplace.cpp:247-259
    float gamma = computeGamma(1.f, stepX);
    float lambda = 1.f;
    EvalScalars sc = evaluate(u_, nullptr, nullptr, gA_, lambda, gamma);
    checkFinite(sc);
    long evals = 0;
    gamma = computeGamma(static_cast<float>(sc.tau), stepX);
    lambda = sc.sum_den > 0.0 ? static_cast<float>(sc.sum_wl / sc.sum_den) : 1.f;
    if (!std::isfinite(lambda) || lambda <= 0.f) lambda = 1.f;
    float tau = static_cast<float>(sc.tau);
    double hpwl = sc.hpwl;
    res->initial_hpwl = hpwl;
    res->initial_overflow = tau;
    res->initial_lambda = lambda;
    const float targetTau = stage == 2 ? cfg_.cgp_target_overflow : cfg_.target_overflow;
    const int maxIter = stage == 1 ? cfg_.filler_only_iters : cfg_.gp_max_iter;
    // hpwlReference (src/gp/lambda_schedule.h:11-20): a zero HPWL under the
    // relative schedule takes delta_hpwl_ref as it is; otherwise the scale must
    // be positive and finite.
    const bool relSched = cfg_.lambda_relative && stage != 1;
    double effRef = relSched ? hpwl * cfg_.delta_hpwl_relative
                             : static_cast<double>(cfg_.delta_hpwl_ref);
    if (relSched && hpwl == 0.0) effRef = cfg_.delta_hpwl_ref;
    else if (!(effRef > 0.0) || !std::isfinite(effRef))
        throw InvalidArg("invalid effective HPWL reference");

    auto deliver = [&](const T* pos) {
        scatterPos(pos);
        double hp = 0;
        res->overflow = static_cast<float>(deliveredMetrics(&hp));
        res->hpwl = hp;
    };
    if (stage != 1 && tau < targetTau) {   // eplace.cpp:278-296: no step at all
        res->converged = 1;
        deliver(u_);
        return;
    }

    // Nesterov init (nesterov.cpp:171-183) and first step size (:190-218)
    const int maxBt = cfg_.max_backtrack > 0 ? cfg_.max_backtrack : 10;
    SPHIP_CHECK(hipMemcpyAsync(v_, u_, bytes2n, hipMemcpyDeviceToDevice, s_));
    SPHIP_CHECK(hipMemcpyAsync(uPrev_, u_, bytes2n, hipMemcpyDeviceToDevice, s_));
    SPHIP_CHECK(hipMemcpyAsync(vPrev_, u_, bytes2n, hipMemcpyDeviceToDevice, s_));
    SPHIP_CHECK(hipMemsetAsync(g_, 0, bytes2n, s_));
    SPHIP_CHECK(hipMemsetAsync(gPrev_, 0, bytes2n, s_));
    double a = 1.0;   // Nesterov a_k, kept in T
    double alpha0 = 1.0;
    {
        if (curPos_ != u_) scatterPos(u_);
        EvalScalars sA = evaluate(u_, nullptr, nullptr, gA_, lambda, gamma);
        checkFinite(sA);
        const double posNorm = std::sqrt(sA.dx2), gNorm = std::sqrt(sA.gg);
        if (gNorm > 0.0 && std::isfinite(gNorm)) {
            const double scale = 0.1 * (posNorm > 0.0 ? posNorm : 1.0) / gNorm;
            hostTimed(SP_HIP_K_UPDATE, [&] {
                launchUpdate<T>(dd_, u_, gA_, nullptr, static_cast<double>(static_cast<T>(scale)),
                                0.0, probe_, nullptr, nullptr, BufTable<T>{}, s_);
            });
            curPos_ = probe_;
            EvalScalars sB = evaluate(probe_, u_, gA_, gB_, lambda, gamma);
            checkFinite(sB);
            const double dx = std::sqrt(sB.dx2), dg = std::sqrt(sB.dg2);
            if (dx > 0.0 && dg > 0.0 && std::isfinite(dg)) {
                const double step = 1.0 / (dg / dx);
                alpha0 = (std::isfinite(step) && step > 0.0) ? static_cast<double>(static_cast<T>(step))
                                                             : 1.0;
            }
        }
    }

    if (cfg_.loop_mode != 0) {
        runDevice(res, rows, maxRows, lambda, gamma, alpha0, effRef, targetTau, maxIter, maxBt,
                  stepX);
        return;
    }

    struct Snap {
        float tau;
        double hpwl;
        int iter;
        T* buf;
    };
    std::vector<Snap> snaps;
    std::vector<T*> spare(snapBufs_.begin(), snapBufs_.end());
    float bestTau = std::numeric_limits<float>::infinity();
    float refTau = std::numeric_limits<float>::infinity();
    int sinceImprove = 0;
    const float minGain = std::clamp(cfg_.gp_stagnation_min_gain, 0.f, 0.5f);
    double prevHpwl = 0;
    bool havePrev = false;
    double alpha = 1.0;
    double momPrev = 0.0;
    const T* finalPos = nullptr;

    const auto t0 = std::chrono::steady_clock::now();
    for (int iter = 0; iter < maxIter; ++iter) {
        const float evalLambda = lambda, evalGamma = gamma;
        // ---- NesterovOptimizer::step (nesterov.cpp:51-122)
        if (curPos_ != u_) scatterPos(u_);
        sc = evaluate(u_, uPrev_, gPrev_, g_, lambda, gamma);
        checkFinite(sc);
        ++evals;
        if (cfg_.gp_backtrack && iter > 0) {
            for (int tries = 1; tries < maxBt; ++tries) {   // redoIfOvershot, :146-169
                const double dx = std::sqrt(sc.dx2), dg = std::sqrt(sc.dg2);
                if (!(dx > 0.0) || !(dg > 0.0) || !std::isfinite(dg)) break;
                const double lip = dx / dg;
                if (lip > 0.95 * alpha) break;
                alpha = static_cast<double>(static_cast<T>(lip));
                launchUpdate<T>(dd_, uPrev_, gPrev_, vPrev_, alpha, momPrev, v_, u_, nullptr,
                                BufTable<T>{}, s_);
                curPos_ = u_;
                sc = evaluate(u_, uPrev_, gPrev_, g_, lambda, gamma);
                checkFinite(sc);
                ++evals;
            }
        }
        // step size (nesterov.cpp:88, :221-240)
        if (iter == 0) {
            alpha = alpha0;
        } else {
            const double dx = std::sqrt(sc.dx2), dg = std::sqrt(sc.dg2);
            if (dx > 0.0 && dg > 0.0) {
                const double lipStep = dx / dg;
                double next = lipStep;
                if (cfg_.use_bb) {
                    const double yy = dg * dg;
                    next = std::min(dx / dg, lipStep);
                    if (yy > 0.0) {
                        const double bb = sc.sy / yy;
                        if (bb > 0.0 && std::isfinite(bb)) next = bb;
                    }
                }
                alpha = static_cast<double>(static_cast<T>(next));
            }
        }
        const T aT = static_cast<T>(a);
        const T aNext = T(0.5) * (T(1) + std::sqrt(T(4) * aT * aT + T(1)));
        const T mom = (aT - T(1)) / aNext;
        launchUpdate<T>(dd_, u_, g_, v_, alpha, static_cast<double>(mom), vNew_, uNew_, nullptr,
                        BufTable<T>{}, s_);
        // rotate (nesterov.cpp:109-118)
        std::swap(uPrev_, u_);
        std::swap(gPrev_, g_);
        std::swap(v_, vNew_);
        std::swap(vPrev_, vNew_);
        std::swap(u_, uNew_);
        curPos_ = u_;
        momPrev = static_cast<double>(mom);
        a = static_cast<double>(aNext);
        const T* lastEval = uPrev_;   // the point the last evaluation used

        tau = static_cast<float>(sc.tau);
        hpwl = sc.hpwl;
        // schedules (eplace.cpp:371-380)
        const double dH = havePrev ? hpwl - prevHpwl : 0.0;
        const float mu = havePrev ? lambdaMultiplier(dH, effRef) : 1.f;
        if (havePrev) lambda *= mu;
        prevHpwl = hpwl;
        havePrev = true;
        gamma = computeGamma(tau, stepX);
        if (!std::isfinite(lambda) || !std::isfinite(gamma))
            throw NonFinite("non-finite schedule parameters");
        if (rows && iter < maxRows) {
            sp_hip_iter_metrics& m = rows[iter];
            m.iter = iter;
            m.evaluations = static_cast<int32_t>(evals);
            m.hpwl = hpwl;
            m.overflow = tau;
            m.lambda = lambda;
            m.gamma = gamma;
            m.step_size = alpha;
            m.grad_norm_wl = static_cast<float>(sc.sum_wl);
            m.grad_norm_den = static_cast<float>(sc.sum_den);
            m.eval_lambda = evalLambda;
            m.eval_gamma = evalGamma;
            m.delta_hpwl = dH;
            m.lambda_mu = mu;
            m.elapsed_ms = std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - t0).count();
        }
        res->iterations = iter + 1;
        if (stage != 1 && tau < targetTau) {   // deliver the look-ahead point, :412-419
            res->converged = 1;
            finalPos = lastEval;
            break;
        }
        if (stage != 1 && cfg_.gp_stagnation_window > 0) {   // :421-473
            if (tau < refTau * (1.f - minGain)) {
                refTau = tau;
                sinceImprove = 0;
            } else {
                ++sinceImprove;
            }
            bestTau = std::min(bestTau, tau);
            const float tol = bestTau * (1.f + minGain);
            while (!snaps.empty() && snaps.front().tau > tol) {
                spare.push_back(snaps.front().buf);
                snaps.erase(snaps.begin());
            }
            if (snaps.empty() || tau * (1.f + 0.5f * minGain) <= snaps.back().tau) {
                T* buf;
                if (!spare.empty()) {
                    buf = spare.back();
                    spare.pop_back();
                } else {
                    buf = dalloc<T>(2 * static_cast<size_t>(n));
                    snapBufs_.push_back(buf);
                }
                SPHIP_CHECK(hipMemcpyAsync(buf, lastEval, bytes2n, hipMemcpyDeviceToDevice, s_));
                snaps.push_back(Snap{tau, hpwl, iter, buf});
            }
            if (sinceImprove >= cfg_.gp_stagnation_window) {
                const Snap* pick = &snaps.front();
                for (const Snap& sn : snaps)
                    if (sn.hpwl < pick->hpwl) pick = &sn;
                finalPos = pick->buf;
                res->stagnated = 1;
                res->rollback_iter = pick->iter;
                break;
            }
        }
    }
    SPHIP_CHECK(hipStreamSynchronize(s_));
    res->seconds_loop =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->evaluations = evals;
    deliver(finalPos ? finalPos : v_);
    if (!std::isfinite(res->hpwl) || !std::isfinite(res->overflow))
        throw NonFinite("non-finite delivered metrics");
}

template <typename T>
void Engine<T>::launchSlot(bool events, int slot) {
    const T qscale = static_cast<T>(std::ldexp(1.0, fracBits_));
    const int* skip = reinterpret_cast<const int*>(reinterpret_cast<const char*>(dCtl_) +
                                                   offsetof(Ctl, done));
    auto b = [&](int f) { if (events) beginK(f, slot); };
    auto e = [&](int f) { if (events) endK(f, slot); };
    // every sort_period-th slot holds the sort passes; they return at once
    // unless the previous scatter found too many nodes outside their tile
    // window (ctl->sortNext)
    const bool sortSlot = slot % sh_.sort_period == 0;
    if (sortSlot) {
        b(SP_HIP_K_SORT);
        launchTileSort<T>(dd_, geom_, dCtl_, s_);
        e(SP_HIP_K_SORT);
    }
    b(SP_HIP_K_SCATTER);
    launchScatter<T>(dd_, geom_, qscale, nodeQ_, fillQ_, sortSlot, ws_, dCtl_, s_);
    e(SP_HIP_K_SCATTER);
    // On the FFT path the first Poisson pass finalizes the density grids
    // itself (profiled or not, so the tau partials are the same either way;
    // the finalize family then has no launches of its own).
    const DensityIn<T> din{geom_, dd_.tiles, nodeQ_, fillQ_, terminal_, base_,
                           std::ldexp(1.0, -fracBits_), binArea_, rhoShift_, ws_.region(0),
                           nullptr, nullptr};
    const bool ownFinalize = !poisson_->takesDensity();
    if (ownFinalize) {
        b(SP_HIP_K_FINALIZE);
        launchFinalize<T>(geom_, dd_.tiles, nodeQ_, fillQ_, terminal_, base_,
                          std::ldexp(1.0, -fracBits_), binArea_, rhoShift_, poisson_->rho(),
                          nullptr, nullptr, ws_, dCtl_, s_);
        e(SP_HIP_K_FINALIZE);
    }
    // Unprofiled slots run the WA pass's blocks inside the Poisson column
    // pass's launch (same blocks, same partials); profiled slots launch the
    // two apart so that each family has its own events.
    bool fused = false;
    if (!events && fuseWa_) {
        const int waBlocks = dd_.waBlocks;
        int densityBlocks = 0;
        fused = poisson_->solveWithWa(
            WaTask<T>{dd_, T(0), cfg_.ignore_net_degree, ws_.region(1), dCtl_}, waBlocks, skip,
            ownFinalize ? nullptr : &din, &densityBlocks);
        if (fused) {
            ws_.blocks[1] = waBlocks;
            if (!ownFinalize) ws_.blocks[0] = densityBlocks;
        }
    }
    if (!fused) {
        b(SP_HIP_K_POISSON);
        if (ownFinalize) poisson_->solve(false, skip);
        else ws_.blocks[0] = poisson_->solveFromDensity(din, skip);
        e(SP_HIP_K_POISSON);
        b(SP_HIP_K_WA_NETS);
        launchWaNets<T>(dd_, 0.f, cfg_.ignore_net_degree, ws_, dCtl_, s_);
        e(SP_HIP_K_WA_NETS);
    }
    b(SP_HIP_K_NODE_GRAD);
    launchNodeGrad<T>(dd_, geom_, poisson_->fieldXY(), 0.f, 0.f, nullptr,
                      nullptr, nullptr, nullptr, nullptr, nullptr, ws_, dCtl_, bufTab_, s_);
    e(SP_HIP_K_NODE_GRAD);
    b(SP_HIP_K_REDUCE);
    launchReduce<T>(ws_, movableAreaScaled_, true, dCtl_, dRows_, s_);
    e(SP_HIP_K_REDUCE);
    b(SP_HIP_K_UPDATE);
    launchUpdate<T>(dd_, nullptr, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr, dCtl_, bufTab_, s_);
    e(SP_HIP_K_UPDATE);
}

// The stage loop with the decisions on the device (loop_mode 1 and 2).
template <typename T>
void Engine<T>::runDevice(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows,
                          float lambda, float gamma, double alpha0, double effRef, float targetTau,
                          int maxIter, int maxBt, float stepX) {
    static_assert(kNumBufs == 8 + 4, "the table lists the eight roles and four snapshots");
    const BufTable<T> table{{u_, v_, uPrev_, vPrev_, uNew_, vNew_, g_, gPrev_, snapPool_[0],
                             snapPool_[1], snapPool_[2], snapPool_[3]}};
    Ctl c{};
    for (int r = 0; r < B_ROLES; ++r) c.cur[r] = c.nxt[r] = r;   // table order = role order
    c.mode = kModeNone;
    c.maxIter = maxIter;
    c.maxBt = maxBt;
    c.backtrack = cfg_.gp_backtrack;
    c.useBB = cfg_.use_bb;
    c.stage = cfg_.stage;
    c.window = cfg_.gp_stagnation_window;
    c.finalBuf = -1;
    c.rollbackIter = -1;
    c.snapWrite = -1;
    c.snapSrc = -1;
    c.nFree = kMaxSnaps;
    for (int i = 0; i < kMaxSnaps; ++i) c.freeSnap[i] = 8 + kMaxSnaps - 1 - i;
    c.a = 1.0;
    c.alpha = 1.0;
    c.alpha0 = alpha0;
    c.lambda = lambda;
    c.gamma = gamma;
    c.targetTau = targetTau;
    c.stepX = stepX;
    c.minGain = std::clamp(cfg_.gp_stagnation_min_gain, 0.f, 0.5f);
    c.bestTau = std::numeric_limits<float>::infinity();
    c.refTau = std::numeric_limits<float>::infinity();
    c.effRef = effRef;
    c.sortNext = 1;
    c.spillLimit = kSpillFrac * dd_.nDens;
    if (maxIter <= 0) {   // no step at all: deliver the clamped start
        res->iterations = 0;
        scatterPos(u_);
        double hp = 0;
        res->overflow = static_cast<float>(deliveredMetrics(&hp));
        res->hpwl = hp;
        return;
    }
    if (rowsCap_ < maxIter) {
        if (dRows_) SPHIP_CHECK(hipFree(dRows_));
        dRows_ = dalloc<DevRow>(maxIter);
        devBytes_ += (maxIter - rowsCap_) * sizeof(DevRow);
        rowsCap_ = maxIter;
        if (graphExec_) {   // the graph holds the old rows pointer
            SPHIP_CHECK(hipGraphExecDestroy(graphExec_));
            SPHIP_CHECK(hipGraphDestroy(graph_));
            graphExec_ = nullptr;
            graph_ = nullptr;
        }
    }
    SPHIP_CHECK(hipMemcpyAsync(dCtl_, &c, sizeof(Ctl), hipMemcpyHostToDevice, s_));
    // the loop kernels take the table by value: a graph captured with other
    // buffers is dropped
    if (graphExec_ && std::memcmp(&table, &bufTab_, sizeof(table)) != 0) {
        SPHIP_CHECK(hipGraphExecDestroy(graphExec_));
        SPHIP_CHECK(hipGraphDestroy(graph_));
        graphExec_ = nullptr;
        graph_ = nullptr;
    }
    bufTab_ = table;
    if (curPos_ != u_) scatterPos(u_);
    SPHIP_CHECK(hipStreamSynchronize(s_));

    // Profiling launches the slots directly (as loop_mode 1): the profiled
    // kernels carry event pairs, which a captured graph does not time.
    const bool graph = cfg_.loop_mode == 2 && profMask_ == 0;
    const int slots = loopSlots();
    if (profMask_) ensureEvents(slots);
    if (graph && !graphExec_) {
        // capture `slots` iterations
        SPHIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
            for (int i = 0; i < slots; ++i) launchSlot(false, i);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SPHIP_CHECK(hipStreamEndCapture(s_, &graph_));
        SPHIP_CHECK(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0));
    }
    const int evMask = graph ? 0 : profMask_;
    loopGraph_ = graph;
    loopSlots_ = slots;
    loopLaunches_ = 0;
    maxLaunches_ = (static_cast<long>(maxIter) * maxBt + 16) / slots + 2;
    inflight_ = 0;
    ringHead_ = 0;
    loopT0_ = std::chrono::steady_clock::now();
    if (asyncStart_) {   // the caller drives the launches (launchAsync / pollAsync)
        asyncReady_ = true;
        return;
    }
    if (evMask) {
        // profiled: the slots one at a time (the same slot sequence as the
        // graph holds), each slot's events harvested, and no slot launched
        // once the stage has ended, so every timed launch did its work
        bool done = false;
        for (long launch = 0; !done; ++launch) {
            if (launch > maxLaunches_) throw std::runtime_error("device loop did not terminate");
            for (int i = 0; i < slots && !done; ++i) {
                launchSlot(true, i);
                SPHIP_CHECK(hipMemcpyAsync(hFlag_,
                                           reinterpret_cast<char*>(dCtl_) + offsetof(Ctl, done),
                                           sizeof(int), hipMemcpyDeviceToHost, s_));
                SPHIP_CHECK(hipStreamSynchronize(s_));
                harvestK(evMask, i, i + 1);
                done = *hFlag_ != 0;
            }
        }
    } else {
        // kAsyncDepth launches in flight: the next one is queued before the
        // host has looked at the previous one's done flag; launches after
        // the stage ended return at once (every kernel checks ctl->done)
        for (int k = 0; k < kAsyncDepth; ++k) launchAsync();
        while (!pollAsync(true)) launchAsync();
    }
    finishDevice(res, rows, maxRows);
}

template <typename T>
void Engine<T>::launchAsync() {
    if (inflight_ >= kAsyncDepth) throw std::logic_error("device loop: too many launches in flight");
    if (++loopLaunches_ > maxLaunches_) throw std::runtime_error("device loop did not terminate");
    if (loopGraph_) {
        SPHIP_CHECK(hipGraphLaunch(graphExec_, s_));
    } else {
        for (int i = 0; i < loopSlots_; ++i) launchSlot(false, i);
    }
    const int k = (ringHead_ + inflight_) % kAsyncDepth;
    SPHIP_CHECK(hipMemcpyAsync(hFlag_ + k, reinterpret_cast<char*>(dCtl_) + offsetof(Ctl, done),
                               sizeof(int), hipMemcpyDeviceToHost, s_));
    SPHIP_CHECK(hipEventRecord(loopEv_[k], s_));
    ++inflight_;
}

template <typename T>
int Engine<T>::pollAsync(bool wait) {
    if (inflight_ == 0) return 0;
    hipEvent_t ev = loopEv_[ringHead_];
    if (wait) {
        SPHIP_CHECK(hipEventSynchronize(ev));
    } else {
        const hipError_t q = hipEventQuery(ev);
        if (q == hipErrorNotReady) return -1;
        SPHIP_CHECK(q);
    }
    const int done = hFlag_[ringHead_] != 0 ? 1 : 0;
    ringHead_ = (ringHead_ + 1) % kAsyncDepth;
    --inflight_;
    return done;
}

// The end of the device loop: the loop state, the metrics rows and the
// delivered position.
template <typename T>
void Engine<T>::finishDevice(sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int maxRows) {
    SPHIP_CHECK(hipStreamSynchronize(s_));
    inflight_ = 0;
    res->seconds_loop =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - loopT0_).count();
    Ctl c{};
    SPHIP_CHECK(hipMemcpyAsync(&c, dCtl_, sizeof(Ctl), hipMemcpyDeviceToHost, s_));
    SPHIP_CHECK(hipStreamSynchronize(s_));
    if (c.status == kNonFinite)
        throw NonFinite("GP/evaluate: non-finite objective, overflow, force norm or gradient");
    res->iterations = c.iter;
    res->evaluations = c.evals;
    res->converged = c.status == kConverged;
    res->stagnated = c.status == kStagnated;
    res->rollback_iter = c.rollbackIter;
    if (rows && maxRows > 0) {
        const int nr = std::min(c.iter, maxRows);
        std::vector<DevRow> dr(nr);
        SPHIP_CHECK(hipMemcpy(dr.data(), dRows_, nr * sizeof(DevRow), hipMemcpyDeviceToHost));
        int rateKHz = 100000;
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
            rateKHz <= 0)
            rateKHz = 100000;
        for (int i = 0; i < nr; ++i) {
            sp_hip_iter_metrics& m = rows[i];
            const DevRow& r = dr[i];
            m.iter = r.iter;
            m.evaluations = r.evaluations;
            m.hpwl = r.hpwl;
            m.overflow = r.overflow;
            m.lambda = r.lambda;
            m.gamma = r.gamma;
            m.step_size = r.step;
            m.grad_norm_wl = r.gradWl;
            m.grad_norm_den = r.gradDen;
            m.eval_lambda = r.evalLambda;
            m.eval_gamma = r.evalGamma;
            m.delta_hpwl = r.dH;
            m.lambda_mu = r.mu;
            m.elapsed_ms = static_cast<double>(r.ticks - dr[0].ticks) / rateKHz;
        }
    }
    if (c.finalBuf < 0 || c.finalBuf >= kNumBufs) throw std::runtime_error("device loop: no result");
    double hp = 0;
    scatterPos(bufTab_.p[c.finalBuf]);
    res->overflow = static_cast<float>(deliveredMetrics(&hp));
    res->hpwl = hp;
}

// ===========================================================================
class PoissonHandleBase {
public:
    virtual ~PoissonHandleBase() = default;
    virtual void solveHost(const float* rho, float* phi, float* ex, float* ey) = 0;
    virtual void solveDev(const void* rho, void* out3) = 0;
    virtual void solveAsync(const void* rho, void* out3, hipStream_t caller) = 0;
    virtual const char* name() const = 0;
};

template <typename T>
class PoissonHandle final : public PoissonHandleBase {
public:
    PoissonHandle(int nx, int ny, float sx, float sy, unsigned opts = 0) : nx_(nx), ny_(ny) {
        requireDevice();
        if (nx < 2 || ny < 2) throw InvalidArg("Poisson: grid must be at least 2x2");
        if (!(sx > 0.f) || !(sy > 0.f)) throw InvalidArg("Poisson: bin step must be positive");
        SPHIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        try {
            plan_ = std::make_unique<PoissonPlan<T>>(nx, ny, sx, sy, s_, false, nullptr, opts);
            const size_t n = static_cast<size_t>(nx) * ny;
            SPHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hIn_), n * sizeof(T), hipHostMallocDefault));
            SPHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hOut_), 3 * n * sizeof(T), hipHostMallocDefault));
            dIn_ = dalloc<T>(n);
            dOut_ = dalloc<T>(3 * n);
            for (hipEvent_t* e : {&evIn_, &evOut_})
                SPHIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        } catch (...) {
            cleanup();
            throw;
        }
    }
    ~PoissonHandle() override { cleanup(); }

    void solveHost(const float* rho, float* phi, float* ex, float* ey) override {
        if (!rho || !phi || !ex || !ey) throw InvalidArg("Poisson: null span");
        const size_t n = static_cast<size_t>(nx_) * ny_;
        // subtracting the mean changes only the discarded DC term and keeps
        // the transform inputs small
        double mean = 0;
        for (size_t i = 0; i < n; ++i) mean += rho[i];
        mean /= static_cast<double>(n);
        for (size_t i = 0; i < n; ++i) hIn_[i] = static_cast<T>(rho[i] - mean);
        SPHIP_CHECK(hipMemcpyAsync(dIn_, hIn_, n * sizeof(T), hipMemcpyHostToDevice, s_));
        run(dIn_, dOut_);
        SPHIP_CHECK(hipMemcpyAsync(hOut_, dOut_, 3 * n * sizeof(T), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
        for (size_t i = 0; i < n; ++i) {
            phi[i] = static_cast<float>(hOut_[i]);
            ex[i] = static_cast<float>(hOut_[n + i]);
            ey[i] = static_cast<float>(hOut_[2 * n + i]);
        }
    }

    void solveDev(const void* rho, void* out3) override {
        if (!rho || !out3) throw InvalidArg("Poisson: null device pointer");
        run(static_cast<const T*>(rho), static_cast<T*>(out3));
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }

    // Ordered on the caller's stream: the handle's stream waits for the
    // caller's work so far, the caller's stream for the solve.
    void solveAsync(const void* rho, void* out3, hipStream_t caller) override {
        if (!rho || !out3) throw InvalidArg("Poisson: null device pointer");
        SPHIP_CHECK(hipEventRecord(evIn_, caller));
        SPHIP_CHECK(hipStreamWaitEvent(s_, evIn_, 0));
        run(static_cast<const T*>(rho), static_cast<T*>(out3));
        SPHIP_CHECK(hipEventRecord(evOut_, s_));
        SPHIP_CHECK(hipStreamWaitEvent(caller, evOut_, 0));
    }

    const char* name() const override {
        if (plan_->usesFft()) return sizeof(T) == 4 ? "hip-fft" : "hip-fft-fp64";
        if (plan_->usesBf16x3()) return "hip-mfma-bf16x3";
        return sizeof(T) == 4 ? "hip-mfma" : "hip-mfma-fp64";
    }

private:
    void run(const T* rho, T* out3) {
        const size_t n = static_cast<size_t>(nx_) * ny_;
        const size_t ld = plan_->ld();
        SPHIP_CHECK(hipMemcpy2DAsync(plan_->rho(), ld * sizeof(T), rho, ny_ * sizeof(T),
                                     ny_ * sizeof(T), nx_, hipMemcpyDeviceToDevice, s_));
        plan_->solve(true);
        T* outs[3] = {out3, out3 + n, out3 + 2 * n};
        T* srcs[3] = {plan_->phi(), plan_->fieldX(), plan_->fieldY()};
        for (int i = 0; i < 3; ++i)
            SPHIP_CHECK(hipMemcpy2DAsync(outs[i], ny_ * sizeof(T), srcs[i], ld * sizeof(T),
                                         ny_ * sizeof(T), nx_, hipMemcpyDeviceToDevice, s_));
    }
    void cleanup() {
        if (s_) (void)hipStreamSynchronize(s_);
        plan_.reset();
        if (hIn_) (void)hipHostFree(hIn_);
        if (hOut_) (void)hipHostFree(hOut_);
        if (dIn_) (void)hipFree(dIn_);
        if (dOut_) (void)hipFree(dOut_);
        for (hipEvent_t e : {evIn_, evOut_})
            if (e) (void)hipEventDestroy(e);
        evIn_ = evOut_ = nullptr;
        if (s_) (void)hipStreamDestroy(s_);
        hIn_ = hOut_ = dIn_ = dOut_ = nullptr;
        s_ = nullptr;
    }
    int nx_, ny_;
    hipStream_t s_ = nullptr;
    std::unique_ptr<PoissonPlan<T>> plan_;
    T *hIn_ = nullptr, *hOut_ = nullptr, *dIn_ = nullptr, *dOut_ = nullptr;
    hipEvent_t evIn_ = nullptr, evOut_ = nullptr;
};

}  // namespace

struct BatchSet::Impl {
    int slots = 1;
    int device = 0;
    hipStream_t s = nullptr;   // from the stream pool
    std::vector<Engine<float>*> members;
    std::vector<long> launches, bound;   // per member: launches since joining, limit
    std::unique_ptr<DevArena> arena;     // argument tables and block maps of the graph
    int* doneDev = nullptr;              // per member: ctl->done after a launch
    int* doneHost = nullptr;             // pinned, kAsyncDepth rows of doneCap
    int doneCap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[kAsyncDepth] = {};
    int head = 0, inflight = 0;

    void dropGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        arena.reset();
    }
    // waits for every launch in flight
    void drain() {
        if (inflight == 0) return;
        SPHIP_CHECK(hipStreamSynchronize(s));
        inflight = 0;
        head = 0;
    }
    void reserveDone(int m) {
        if (m <= doneCap) return;
        const int cap = std::max(m, 2 * doneCap);
        if (doneDev) (void)hipFree(doneDev);
        if (doneHost) (void)hipHostFree(doneHost);
        doneDev = nullptr;
        doneHost = nullptr;
        doneCap = 0;
        SPHIP_CHECK(hipMalloc(reinterpret_cast<void**>(&doneDev), cap * sizeof(int)));
        SPHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&doneHost),
                                  kAsyncDepth * cap * sizeof(int), hipHostMallocDefault));
        doneCap = cap;
    }
    void capture();
};

BatchSet::BatchSet(int slots) : impl_(std::make_unique<Impl>()) {
    requireDevice();
    impl_->slots = std::max(1, slots);
    SPHIP_CHECK(hipGetDevice(&impl_->device));
    impl_->s = takeStream(impl_->device);
    try {
        for (hipEvent_t& e : impl_->ev) SPHIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    } catch (...) {
        for (hipEvent_t e : impl_->ev)
            if (e) (void)hipEventDestroy(e);
        giveStream(impl_->device, impl_->s);
        throw;
    }
}

BatchSet::~BatchSet() {
    Impl& b = *impl_;
    if (b.s) (void)hipStreamSynchronize(b.s);
    b.dropGraph();
    for (hipEvent_t e : b.ev)
        if (e) (void)hipEventDestroy(e);
    if (b.doneDev) (void)hipFree(b.doneDev);
    if (b.doneHost) (void)hipHostFree(b.doneHost);
    giveStream(b.device, b.s);
}

int BatchSet::inFlight() const { return impl_->inflight; }
int BatchSet::size() const { return static_cast<int>(impl_->members.size()); }

void BatchSet::assign(const std::vector<EngineBase*>& members) {
    Impl& b = *impl_;
    b.drain();
    b.dropGraph();
    b.members.clear();
    b.launches.clear();
    b.bound.clear();
    for (EngineBase* e : members) {
        auto* f = dynamic_cast<Engine<float>*>(e);
        if (!f) throw InvalidArg("batched runs take FP32 instances");
        b.members.push_back(f);
        b.launches.push_back(0);
        b.bound.push_back(f->maxLoopLaunches(b.slots));
    }
    if (b.members.empty()) return;
    b.reserveDone(static_cast<int>(b.members.size()));
    b.capture();
}

// Block maps and argument tables for the members, then the graph of
// `slots` batched slots.
void BatchSet::Impl::capture() {
    const int m = static_cast<int>(members.size());
    std::vector<SlotArgs<float>> args(m);
    std::vector<PoissonBatchArgs<float>> pargs(m);
    std::vector<std::vector<int4>> maps(BK_KINDS);
    std::vector<int> dims, lds;
    std::vector<std::array<std::vector<int4>, 3>> fftMaps;   // per grid size, per pass
    // the sort passes at the shortest period among the members (the sort
    // only reorders records, so a member's results do not depend on it)
    int sortPeriod = kSortPeriod;
    for (int j = 0; j < m; ++j) sortPeriod = j == 0 ? members[j]->sortPeriod()
                                                    : std::min(sortPeriod, members[j]->sortPeriod());
    for (int j = 0; j < m; ++j) {
        int counts[BK_KINDS];
        members[j]->batchShare(&args[j], &pargs[j], counts, doneDev);
        for (int k = 0; k < BK_KINDS; ++k)
            for (int q = 0; q < counts[k]; ++q) maps[k].push_back(int4{j, q, counts[k], 0});
        const int dim = members[j]->gridDim();
        const int gi = static_cast<int>(std::find(dims.begin(), dims.end(), dim) - dims.begin());
        if (gi == static_cast<int>(dims.size())) {
            dims.push_back(dim);
            lds.push_back(members[j]->gridLd());   // one padded stride per grid size
            fftMaps.emplace_back();
        }
        for (int pass = 0; pass < 3; ++pass) {
            const int c = poissonBatchBlocks(pass, dim, dim);
            for (int q = 0; q < c; ++q) fftMaps[gi][pass].push_back(int4{j, q, c, 0});
        }
    }
    arena = std::make_unique<DevArena>();
    SlotArgs<float>* dArgs = nullptr;
    PoissonBatchArgs<float>* dPargs = nullptr;
    arena->upload(&dArgs, args);
    arena->upload(&dPargs, pargs);
    std::vector<int4*> dMaps(BK_KINDS, nullptr);
    for (int k = 0; k < BK_KINDS; ++k) arena->upload(&dMaps[k], maps[k]);
    std::vector<std::array<int4*, 3>> dFft(dims.size());
    for (size_t g = 0; g < dims.size(); ++g)
        for (int pass = 0; pass < 3; ++pass) arena->upload(&dFft[g][pass], fftMaps[g][pass]);
    arena->place(s);
    SPHIP_CHECK(hipMemsetAsync(doneDev, 0, m * sizeof(int), s));
    SPHIP_CHECK(hipStreamSynchronize(s));
    auto count = [&](int k) { return static_cast<int>(maps[k].size()); };
    SPHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        for (int slot = 0; slot < slots; ++slot) {
            const int sorted = slot % sortPeriod == 0 ? 1 : 0;
            if (sorted)
                for (int k : {BK_SORT_COUNT, BK_SORT_PLACE})
                    launchSlotBatch<float>(k, dArgs, dMaps[k], count(k), 1, s);
            for (int k : {BK_SCATTER, BK_FINALIZE})
                launchSlotBatch<float>(k, dArgs, dMaps[k], count(k), sorted, s);
            for (size_t g = 0; g < dims.size(); ++g)
                for (int pass = 0; pass < 3; ++pass)
                    launchPoissonBatch<float>(pass, dPargs, dFft[g][pass],
                                              static_cast<int>(fftMaps[g][pass].size()), dims[g],
                                              dims[g], lds[g], s);
            for (int k : {BK_WA, BK_NODE, BK_REDUCE, BK_UPDATE})
                launchSlotBatch<float>(k, dArgs, dMaps[k], count(k), sorted, s);
        }
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    SPHIP_CHECK(hipStreamEndCapture(s, &graph));
    SPHIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
}

void BatchSet::launch() {
    Impl& b = *impl_;
    if (b.members.empty()) throw InvalidArg("batched launch: no members");
    if (b.inflight >= kAsyncDepth) throw InvalidArg("batched launch: ring full");
    const int m = static_cast<int>(b.members.size());
    const int k = (b.head + b.inflight) % kAsyncDepth;
    SPHIP_CHECK(hipGraphLaunch(b.exec, b.s));
    SPHIP_CHECK(hipMemcpyAsync(b.doneHost + static_cast<size_t>(k) * b.doneCap, b.doneDev,
                               m * sizeof(int), hipMemcpyDeviceToHost, b.s));
    SPHIP_CHECK(hipEventRecord(b.ev[k], b.s));
    for (long& l : b.launches) ++l;
    ++b.inflight;
}

int BatchSet::poll(bool wait, std::vector<int>* state) {
    Impl& b = *impl_;
    if (b.inflight == 0) throw InvalidArg("batched poll: nothing in flight");
    const int k = b.head;
    if (wait) {
        SPHIP_CHECK(hipEventSynchronize(b.ev[k]));
    } else {
        const hipError_t q = hipEventQuery(b.ev[k]);
        if (q == hipErrorNotReady) return -1;
        SPHIP_CHECK(q);
    }
    b.head = (b.head + 1) % kAsyncDepth;
    --b.inflight;
    const int m = static_cast<int>(b.members.size());
    // launches the member had when this launch ended
    const long behind = b.inflight;
    state->assign(m, 0);
    for (int j = 0; j < m; ++j) {
        if (b.doneHost[static_cast<size_t>(k) * b.doneCap + j] != 0) (*state)[j] = 1;
        else if (b.launches[j] - behind >= b.bound[j]) (*state)[j] = 2;
    }
    return 0;
}

// Many FP32 stages advanced together by batched launches until every
// instance's stage has ended; each instance's result equals its own run's.
void runBatch(const std::vector<EngineBase*>& engines, const std::vector<char>& looping,
              sp_hip_gp_result* res, int slots) {
    std::vector<EngineBase*> live;
    std::vector<int> liveAt;
    for (size_t i = 0; i < engines.size(); ++i) {
        if (!dynamic_cast<Engine<float>*>(engines[i]))
            throw InvalidArg("batched runs take FP32 instances");
        if (!looping[i]) continue;   // the stage ended before its loop
        live.push_back(engines[i]);
        liveAt.push_back(static_cast<int>(i));
    }
    if (live.empty()) return;
    BatchSet set(slots);
    set.assign(live);
    for (int k = 0; k < kAsyncDepth; ++k) set.launch();
    std::vector<int> state;
    for (;;) {
        set.poll(true, &state);
        bool all = true;
        for (int st : state) {
            if (st == 2) throw std::runtime_error("batched loop did not terminate");
            all = all && st == 1;
        }
        if (all) break;
        set.launch();
    }
    set.assign({});   // waits for the launch still in flight
    for (size_t j = 0; j < live.size(); ++j) live[j]->finishAsync(&res[liveAt[j]]);
}

std::unique_ptr<EngineBase> makeEngine(const sp_hip_design& d, const sp_hip_gp_config& cfg) {
    if (cfg.precision != 32 && cfg.precision != 64)
        throw InvalidArg("precision must be 32 or 64");
    if (cfg.stage < 0 || cfg.stage > 2) throw InvalidArg("stage must be 0, 1 or 2");
    if (!(cfg.target_density > 0.f)) throw InvalidArg("target_density must be positive");
    if (cfg.precision == 64) return std::make_unique<Engine<double>>(d, cfg);
    return std::make_unique<Engine<float>>(d, cfg);
}

size_t footprintEstimate(const sp_hip_design& d, const sp_hip_gp_config& cfg) {
    validateDesign(d);
    if (cfg.precision != 32 && cfg.precision != 64)
        throw InvalidArg("precision must be 32 or 64");
    if (!(cfg.target_density > 0.f)) throw InvalidArg("target_density must be positive");
    const size_t t = cfg.precision == 64 ? 8 : 4;
    const size_t total = static_cast<size_t>(d.num_nodes) + d.num_fillers;
    const size_t pins = d.num_pins, nets = d.num_nets;
    const size_t dim = binDim(d, cfg), bins = dim * dim;
    const size_t padded = static_cast<size_t>(roundUp(static_cast<int>(dim), 64));
    const size_t tiles = (dim + kTileBins - 1) / kTileBins;
    const int maxIter = cfg.stage == 1 ? cfg.filler_only_iters : cfg.gp_max_iter;
    size_t b = 0;
    // per node: coordinates, sizes, flags, index, density records, CSR start
    b += total * (12 * t + 25);
    // per optimised node (at most every node): preconditioner, the state
    // vectors and the snapshot pool
    b += total * (12 + (2 * 11 + 2 * kMaxSnaps) * t);
    // per pin: node, offsets and node-order entry in lane order (at most two
    // lanes per pin: a net of degree d takes fewer than 2d lanes) or as a
    // tail slot, and its gradient; per net
    b += pins * (2 * (8 + 2 * t) + 2 * t) + 4 * (nets + 1);
    // per bin: fixed-point grids, static densities, tile staging, and the
    // transform buffers of the padded grid
    b += bins * 32 + tiles * tiles * 2 * kTileWin * kTileWin * 8 + padded * padded * 8 * t;
    b += 4 * padded * 2 * t * 2 + 4 * padded * t;   // twiddle tables, u and v
    if (!PoissonPlan<float>::fftShape(static_cast<int>(dim)))
        b += 6 * padded * padded * t;   // the dense transform matrices
    // reduction partials (4 regions of at most one partial per 64 items),
    // tile counters, metrics rows
    b += 4 * 8 * 8 * ((std::max({total, pins, bins}) + 63) / 64 + 2) + 5 * 4 * (kTileKeys + 1);
    b += static_cast<size_t>(std::max(maxIter, 0)) * sizeof(DevRow);
    return b + 65536;
}

}  // namespace sphip

struct sp_hip_poisson {
    std::unique_ptr<sphip::PoissonHandleBase> impl;
};
struct sp_hip_gp {
    std::unique_ptr<sphip::EngineBase> impl;
};

using sphip::guarded;

extern "C" {

const char* sp_hip_last_error(void) { return sphip::g_lastError.c_str(); }

int sp_hip_device_count(int* count) {
    return guarded([&] {
        if (!count) throw sphip::InvalidArg("null count");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = n;
    });
}

int sp_hip_poisson_create(int nx, int ny, float step_x, float step_y, int flags,
                          sp_hip_poisson** out) {
    return guarded([&] {
        if (!out) throw sphip::InvalidArg("null output handle");
        *out = nullptr;
        auto h = std::make_unique<sp_hip_poisson>();
        if ((flags & SP_HIP_FLAG_FP64) && (flags & SP_HIP_FLAG_BF16X3))
            throw sphip::InvalidArg("Poisson: split-BF16 products are FP32 only");
        const unsigned opts = ((flags & SP_HIP_FLAG_DENSE) ? sphip::kPlanDense : 0u) |
                              ((flags & SP_HIP_FLAG_BF16X3) ? sphip::kPlanBf16x3 : 0u);
        if (flags & SP_HIP_FLAG_FP64)
            h->impl = std::make_unique<sphip::PoissonHandle<double>>(nx, ny, step_x, step_y, opts);
        else
            h->impl = std::make_unique<sphip::PoissonHandle<float>>(nx, ny, step_x, step_y, opts);
        *out = h.release();
    });
}

int sp_hip_poisson_solve(sp_hip_poisson* p, const float* density, float* phi, float* field_x,
                         float* field_y) {
    return guarded([&] {
        if (!p) throw sphip::InvalidArg("null handle");
        p->impl->solveHost(density, phi, field_x, field_y);
    });
}

int sp_hip_poisson_solve_dev(sp_hip_poisson* p, const void* density_dev, void* out3_dev) {
    return guarded([&] {
        if (!p) throw sphip::InvalidArg("null handle");
        p->impl->solveDev(density_dev, out3_dev);
    });
}

int sp_hip_shapes_set(int bin_dim, const sp_hip_shapes* shapes) {
    return guarded([&] {
        if (!shapes) throw sphip::InvalidArg("null shapes");
        if (bin_dim < 0) throw sphip::InvalidArg("shapes: negative bin dimension");
        sphip::validateShapes(*shapes);
        auto& t = sphip::shapeTable();
        std::lock_guard<std::mutex> lock(t.mu);
        for (auto& e : t.entries)
            if (e.dim == bin_dim) {
                e.shapes = *shapes;
                return;
            }
        t.entries.push_back({bin_dim, *shapes});
    });
}

int sp_hip_shapes_get(int bin_dim, sp_hip_shapes* shapes) {
    return guarded([&] {
        if (!shapes) throw sphip::InvalidArg("null shapes");
        *shapes = sphip::shapesFor(bin_dim);
    });
}

int sp_hip_shapes_reset(void) {
    return guarded([&] {
        auto& t = sphip::shapeTable();
        std::lock_guard<std::mutex> lock(t.mu);
        t.entries.clear();
    });
}

int sp_hip_poisson_solve_async(sp_hip_poisson* p, const void* density_dev, void* out3_dev,
                               void* stream) {
    return guarded([&] {
        if (!p) throw sphip::InvalidArg("null handle");
        p->impl->solveAsync(density_dev, out3_dev, static_cast<hipStream_t>(stream));
    });
}

const char* sp_hip_poisson_name(const sp_hip_poisson* p) { return p ? p->impl->name() : "none"; }

void sp_hip_poisson_destroy(sp_hip_poisson* p) { delete p; }

int sp_hip_gp_create(const sp_hip_design* d, const sp_hip_gp_config* cfg, sp_hip_gp** out) {
    return guarded([&] {
        if (!d || !cfg || !out) throw sphip::InvalidArg("null argument");
        *out = nullptr;
        auto h = std::make_unique<sp_hip_gp>();
        h->impl = sphip::makeEngine(*d, *cfg);
        *out = h.release();
    });
}

int sp_hip_gp_info_get(const sp_hip_gp* gp, sp_hip_gp_info* info) {
    return guarded([&] {
        if (!gp || !info) throw sphip::InvalidArg("null argument");
        gp->impl->info(info);
    });
}

int sp_hip_gp_opt_index(const sp_hip_gp* gp, int32_t* idx) {
    return guarded([&] {
        if (!gp || !idx) throw sphip::InvalidArg("null argument");
        gp->impl->optIndex(idx);
    });
}

int sp_hip_gp_eval(sp_hip_gp* gp, const float* x, const float* y, float lambda, float gamma,
                   sp_hip_eval_scalars* out, sp_hip_eval_debug* dbg) {
    return guarded([&] {
        if (!gp || !x || !y || !out) throw sphip::InvalidArg("null argument");
        gp->impl->evalHost(x, y, lambda, gamma, out, dbg);
    });
}

int sp_hip_gp_run(sp_hip_gp* gp, sp_hip_gp_result* res, sp_hip_iter_metrics* rows, int max_rows) {
    return guarded([&] {
        if (!gp || !res) throw sphip::InvalidArg("null argument");
        gp->impl->run(res, rows, rows ? max_rows : 0);
    });
}

int sp_hip_gp_run_batch(int count, const sp_hip_design* designs, const sp_hip_gp_config* cfgs,
                        int slots, sp_hip_gp_result* results, float* const* x, float* const* y) {
    return guarded([&] {
        if (count < 0 || (count > 0 && (!designs || !cfgs || !results)))
            throw sphip::InvalidArg("batch: null argument");
        for (int i = 0; i < count; ++i)
            if (cfgs[i].precision != 32) throw sphip::InvalidArg("batch: instances must be FP32");
        // set-ups (design upload, the stage's first evaluations) on host
        // threads, each instance on its own stream
        std::vector<std::unique_ptr<sphip::EngineBase>> engines(count);
        std::vector<char> looping(count, 0);
        std::vector<std::exception_ptr> errs(count);
        std::atomic<int> next{0};
        auto work = [&] {
            for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
                try {
                    sp_hip_gp_config c = cfgs[i];
                    c.loop_mode = 1;   // the batch drives the loop; no per-instance graph
                    engines[i] = sphip::makeEngine(designs[i], c);
                    looping[i] = engines[i]->begin(&results[i]) ? 0 : 1;
                } catch (...) {
                    errs[i] = std::current_exception();
                }
            }
        };
        const int threads = std::max(1, std::min(8, count));
        std::vector<std::thread> pool;
        for (int t = 1; t < threads; ++t) pool.emplace_back(work);
        work();
        for (auto& t : pool) t.join();
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
        std::vector<sphip::EngineBase*> view;
        for (auto& e : engines) view.push_back(e.get());
        sphip::runBatch(view, looping, results, slots > 0 ? slots : 8);
        for (int i = 0; i < count; ++i)
            if (x && y && x[i] && y[i]) engines[i]->positions(x[i], y[i]);
    });
}

int sp_hip_gp_get_positions(const sp_hip_gp* gp, float* x, float* y) {
    return guarded([&] {
        if (!gp || !x || !y) throw sphip::InvalidArg("null argument");
        gp->impl->positions(x, y);
    });
}

int sp_hip_gp_set_profiling(sp_hip_gp* gp, int on) {
    return guarded([&] {
        if (!gp) throw sphip::InvalidArg("null argument");
        gp->impl->setProfiling(on);
    });
}

int sp_hip_gp_kernel_stats(const sp_hip_gp* gp, double* ms, int64_t* launches, double* bytes,
                           double* flops) {
    return guarded([&] {
        if (!gp) throw sphip::InvalidArg("null argument");
        gp->impl->kernelStats(ms, launches, bytes, flops);
    });
}

void sp_hip_gp_destroy(sp_hip_gp* gp) { delete gp; }

int sp_hip_gemm(int flags, int batch, const void* const* a, const void* const* b, void* const* c,
                int m, int n, int k) {
    return guarded([&] {
        sphip::requireDevice();
        if (!a || !b || !c || batch < 1 || batch > 4) throw sphip::InvalidArg("bad gemm batch");
        if (m <= 0 || n <= 0 || k <= 0 || m % 64 || n % 64 || k % 64)
            throw sphip::InvalidArg("gemm dims must be positive multiples of 64");
        if ((flags & SP_HIP_FLAG_FP64) && (flags & SP_HIP_FLAG_BF16X3))
            throw sphip::InvalidArg("gemm: split-BF16 products are FP32 only");
        if (flags & SP_HIP_FLAG_FP64)
            sphip::launchGemm<double>(batch, reinterpret_cast<const double* const*>(a),
                                      reinterpret_cast<const double* const*>(b),
                                      reinterpret_cast<double* const*>(c), m, n, k, k, k, n, nullptr);
        else
            sphip::launchGemm<float>(batch, reinterpret_cast<const float* const*>(a),
                                     reinterpret_cast<const float* const*>(b),
                                     reinterpret_cast<float* const*>(c), m, n, k, k, k, n, nullptr,
                                     (flags & SP_HIP_FLAG_BF16X3) != 0);
        SPHIP_CHECK(hipDeviceSynchronize());
    });
}

}  // extern "C"
