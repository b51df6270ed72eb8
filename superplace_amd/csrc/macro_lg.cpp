// Macro legalization by simulated annealing with greedy repairs, as the
// reference runs it before Abacus (legalizeMacros, src/lg/macro_sa.cpp:493-1272;
// contact candidates, src/lg/repair_candidates.h:14-45).  Host code: one
// sequential Markov chain followed by sequential repair passes.
//
// The result must equal the reference's bit for bit on the same input and
// seed, so its steps, their order, its random draws (std::mt19937 with
// libstdc++'s uniform distributions, drawn in the same order) and its
// float/double mix are kept:
//  * macros = movable, macro-flagged, not fixed nodes; they are snapped to
//    the site / row grid and clamped into the core first;
//  * cost = 8 overlap / macro area + wirelength / initial wirelength
//    + delta * displacement / (typical macro side * macros), wirelength over
//    the nets with macro pins, kept incrementally (fixed per-net box of the
//    other pins plus each macro's pin box on the net, mirrored when flipped);
//  * overlap (macro-macro and macro-fixed blockage) through a uniform grid
//    index whose candidates are sorted before summing;
//  * the start temperature makes the median worsening of 2000 probe moves
//    accepted with probability init_accept; the temperature falls
//    geometrically to temp_ratio; 70% of moves push an overlapping macro off
//    its worst neighbour, the rest shift, swap or flip a random macro; the
//    best state seen is restored at the end;
//  * repairs, smallest moves first: minimum separation (both tie rules, the
//    better kept), chain pushes, a grid-and-contact search limited to 3 macro
//    sides, again separation and chains, then an unlimited grid search if
//    overlap remains; a last pull of the most displaced macro toward its
//    start when legal and not worse in wirelength;
//  * macros end fixed, flipped ones with mirrored pin x offsets and N<->FN,
//    S<->FS orientation; per-step summaries go to a metrics sink.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <limits>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "abi.h"

namespace sphip {
namespace {

struct Box {
    float lx, ly, hx, hy;
};

inline float overlapOf(const Box& a, const Box& b) {
    const float dx = std::min(a.hx, b.hx) - std::max(a.lx, b.lx);
    const float dy = std::min(a.hy, b.hy) - std::max(a.ly, b.ly);
    return (dx > 0.f && dy > 0.f) ? dx * dy : 0.f;
}

// The design arrays the pass reads, with the caller's in/out positions and
// pin x offsets.
struct View {
    const sp_hip_design& d;
    const float* netW;
    float* x;
    float* y;
    float* offx;
    float pinX(int p) const { return x[d.pin2node[p]] + offx[p]; }
    float pinY(int p) const { return y[d.pin2node[p]] + d.pin_offset_y[p]; }
    bool flag(int i, uint8_t f, const uint8_t* flags) const { return (flags[i] & f) != 0; }
};

// computeHPWL (src/db/hpwl.cpp:13-39) with its fixed 64-chunk summation
// (src/numeric/reduction.h:24-71).
double hpwlOf(const View& v) {
    const int64_t n = v.d.num_nets;
    if (n <= 0) return 0.0;
    double part[64];
    for (int c = 0; c < 64; ++c) {
        const int64_t lo = n * c / 64, hi = n * (c + 1) / 64;
        double s = 0.0;
        for (int64_t k = lo; k < hi; ++k) {
            const int b = v.d.net2pin_start[k], e = v.d.net2pin_start[k + 1];
            if (e - b < 2) continue;
            float lox = std::numeric_limits<float>::max(), hix = -lox, loy = lox, hiy = -lox;
            for (int t = b; t < e; ++t) {
                const int p = v.d.flat_net2pin[t];
                const float px = v.pinX(p), py = v.pinY(p);
                lox = std::min(lox, px);
                hix = std::max(hix, px);
                loy = std::min(loy, py);
                hiy = std::max(hiy, py);
            }
            s += static_cast<double>(hix - lox) + static_cast<double>(hiy - loy);
        }
        part[c] = s;
    }
    double total = 0.0;
    for (double p : part) total += p;
    return total;
}

// Wirelength of the nets with macro pins, updated per moved macro.
class MacroWirelength {
public:
    void build(const View& v, const std::vector<int>& macros) {
        const sp_hip_design& d = v.d;
        const int nets = d.num_nets;
        std::vector<uint8_t> macroPin(d.num_pins, 0);
        for (int m : macros)
            for (int k = d.node2pin_start[m]; k < d.node2pin_start[m + 1]; ++k)
                macroPin[d.flat_node2pin[k]] = 1;
        const float inf = std::numeric_limits<float>::infinity();
        fixedBox_.assign(nets, Box{inf, inf, -inf, -inf});
        hasMacro_.assign(nets, 0);
        for (int n = 0; n < nets; ++n) {
            Box& b = fixedBox_[n];
            for (int k = d.net2pin_start[n]; k < d.net2pin_start[n + 1]; ++k) {
                const int p = d.flat_net2pin[k];
                if (macroPin[p]) {
                    hasMacro_[n] = 1;
                    continue;
                }
                const float px = v.pinX(p), py = v.pinY(p);
                b.lx = std::min(b.lx, px);
                b.hx = std::max(b.hx, px);
                b.ly = std::min(b.ly, py);
                b.hy = std::max(b.hy, py);
            }
        }
        // per macro, its pin box on each of its nets (nets in first-pin order)
        start_.assign(macros.size() + 1, 0);
        std::vector<int> lastMacro(nets, -1), at(nets, -1);
        for (size_t s = 0; s < macros.size(); ++s) {
            const int m = macros[s];
            for (int k = d.node2pin_start[m]; k < d.node2pin_start[m + 1]; ++k) {
                const int p = d.flat_node2pin[k];
                const int n = d.pin2net[p];
                const float ox = v.offx[p], oy = d.pin_offset_y[p];
                if (lastMacro[n] != static_cast<int>(s)) {
                    lastMacro[n] = static_cast<int>(s);
                    at[n] = static_cast<int>(boxes_.size());
                    boxes_.push_back({n, Box{ox, oy, ox, oy}});
                } else {
                    Box& b = boxes_[at[n]].second;
                    b.lx = std::min(b.lx, ox);
                    b.hx = std::max(b.hx, ox);
                    b.ly = std::min(b.ly, oy);
                    b.hy = std::max(b.hy, oy);
                }
            }
            start_[s + 1] = static_cast<int>(boxes_.size());
        }
        // per net, the (box, macro) pairs on it, macros in order
        netStart_.assign(nets + 1, 0);
        for (const auto& nb : boxes_) ++netStart_[nb.first + 1];
        for (int n = 0; n < nets; ++n) netStart_[n + 1] += netStart_[n];
        netBox_.assign(netStart_[nets], 0);
        netSlot_.assign(netStart_[nets], 0);
        std::vector<int> fill(netStart_.begin(), netStart_.end() - 1);
        for (size_t s = 0; s < macros.size(); ++s)
            for (int j = start_[s]; j < start_[s + 1]; ++j) {
                const int at2 = fill[boxes_[j].first]++;
                netBox_[at2] = j;
                netSlot_[at2] = static_cast<int>(s);
            }
        weight_.assign(nets, 1.f);
        if (v.netW)
            for (int n = 0; n < nets; ++n) weight_[n] = v.netW[n];
        cache_.assign(nets, 0.0);
    }

    double total(const std::vector<float>& cx, const std::vector<float>& cy,
                 const std::vector<uint8_t>& flip) {
        double sum = 0.0;
        for (size_t n = 0; n < hasMacro_.size(); ++n) {
            if (!hasMacro_[n]) continue;
            cache_[n] = net(static_cast<int>(n), cx, cy, flip);
            sum += cache_[n];
        }
        return sum;
    }

    double delta(int s, const std::vector<float>& cx, const std::vector<float>& cy,
                 const std::vector<uint8_t>& flip, std::vector<int>& nets,
                 std::vector<double>& vals) const {
        nets.clear();
        vals.clear();
        double dsum = 0.0;
        for (int j = start_[s]; j < start_[s + 1]; ++j) {
            const int n = boxes_[j].first;
            const double val = net(n, cx, cy, flip);
            nets.push_back(n);
            vals.push_back(val);
            dsum += val - cache_[n];
        }
        return dsum;
    }

    void commit(const std::vector<int>& nets, const std::vector<double>& vals) {
        for (size_t i = 0; i < nets.size(); ++i) cache_[nets[i]] = vals[i];
    }

private:
    double net(int n, const std::vector<float>& cx, const std::vector<float>& cy,
               const std::vector<uint8_t>& flip) const {
        Box b = fixedBox_[n];
        for (int t = netStart_[n]; t < netStart_[n + 1]; ++t) {
            const Box& lb = boxes_[netBox_[t]].second;
            const int s = netSlot_[t];
            const float px = cx[s], py = cy[s];
            const bool f = flip[s] != 0;
            const float o0 = f ? -lb.hx : lb.lx, o1 = f ? -lb.lx : lb.hx;
            b.lx = std::min(b.lx, px + o0);
            b.hx = std::max(b.hx, px + o1);
            b.ly = std::min(b.ly, py + lb.ly);
            b.hy = std::max(b.hy, py + lb.hy);
        }
        if (!(b.hx >= b.lx) || !(b.hy >= b.ly)) return 0.0;
        return static_cast<double>(weight_[n]) *
               (static_cast<double>(b.hx - b.lx) + static_cast<double>(b.hy - b.ly));
    }

    std::vector<Box> fixedBox_;
    std::vector<uint8_t> hasMacro_;
    std::vector<std::pair<int, Box>> boxes_;
    std::vector<int> start_, netStart_, netBox_, netSlot_;
    std::vector<float> weight_;
    std::vector<double> cache_;
};

// Macro-macro and macro-blockage overlap through a uniform grid index; the
// only writer of macro positions.
class OverlapIndex {
public:
    void build(const View& v, const uint8_t* flags, const std::vector<int>& macros,
               const std::vector<float>& px, const std::vector<float>& py) {
        const sp_hip_design& d = v.d;
        const size_t n = macros.size();
        w_.resize(n);
        h_.resize(n);
        cx_ = px;
        cy_ = py;
        for (size_t s = 0; s < n; ++s) {
            w_[s] = d.node_w[macros[s]];
            h_[s] = d.node_h[macros[s]];
        }
        for (int i = d.num_movable; i < d.num_nodes; ++i) {
            if (flags[i] & F_NI) continue;
            if (d.node_w[i] <= 0.f || d.node_h[i] <= 0.f) continue;
            const float hw = 0.5f * d.node_w[i], hh = 0.5f * d.node_h[i];
            blocks_.push_back(Box{v.x[i] - hw, v.y[i] - hh, v.x[i] + hw, v.y[i] + hh});
        }
        const double target = std::sqrt(2.0 * static_cast<double>(std::max<size_t>(n, 1)));
        dim_ = std::min(128, std::max(4, static_cast<int>(target + 0.5)));
        ox_ = d.core_lx;
        oy_ = d.core_ly;
        cw_ = std::max(1e-3f, (d.core_hx - d.core_lx) / static_cast<float>(dim_));
        ch_ = std::max(1e-3f, (d.core_hy - d.core_ly) / static_cast<float>(dim_));
        cells_.assign(static_cast<size_t>(dim_) * dim_, {});
        blockCells_.assign(static_cast<size_t>(dim_) * dim_, {});
        for (size_t s = 0; s < n; ++s) insert(static_cast<int>(s));
        for (size_t b = 0; b < blocks_.size(); ++b) {
            int x0, x1, y0, y1;
            range(blocks_[b], x0, x1, y0, y1);
            for (int gx = x0; gx <= x1; ++gx)
                for (int gy = y0; gy <= y1; ++gy) blockCells_[idx(gx, gy)].push_back(static_cast<int>(b));
        }
        own_ = scratch();
    }

    // Query scratch of at(): candidate lists and visit stamps.  at() uses the
    // index's own; a thread querying beside others passes its own to atWith().
    struct Scratch {
        std::vector<int> seen, blockSeen, cand, bcand;
        int tick = 0;
    };
    Scratch scratch() const {
        Scratch sc;
        sc.seen.assign(w_.size(), -1);
        sc.blockSeen.assign(blocks_.size(), -1);
        return sc;
    }

    const std::vector<float>& xs() const { return cx_; }
    const std::vector<float>& ys() const { return cy_; }
    const std::vector<Box>& blocks() const { return blocks_; }
    float width(int s) const { return w_[s]; }
    float height(int s) const { return h_[s]; }
    int dim() const { return dim_; }

    void move(int s, float px, float py) {
        erase(s);
        cx_[s] = px;
        cy_[s] = py;
        insert(s);
    }

    Box rect(int s, float px, float py) const {
        const float hw = 0.5f * w_[s], hh = 0.5f * h_[s];
        return Box{px - hw, py - hh, px + hw, py + hh};
    }

    // Overlap of macro s placed at (px, py) with the other macros and the
    // blockages.
    double at(int s, float px, float py) const { return atWith(own_, s, px, py); }
    double atWith(Scratch& sc, int s, float px, float py) const {
        const Box a = rect(s, px, py);
        int x0, x1, y0, y1;
        range(a, x0, x1, y0, y1);
        gatherMacros(sc, x0, x1, y0, y1, s);
        double sum = 0.0;
        for (int t : sc.cand) sum += overlapOf(a, rect(t, cx_[t], cy_[t]));
        if (!blocks_.empty()) {
            gatherBlocks(sc, x0, x1, y0, y1);
            for (int b : sc.bcand) sum += overlapOf(a, blocks_[b]);
        }
        return sum;
    }

    // The macro or blockage overlapping (s at (px, py)) most: macros first,
    // each kind in index order, strictly larger replaces.
    bool worst(int s, float px, float py, Box& out) const {
        const Box a = rect(s, px, py);
        int x0, x1, y0, y1;
        range(a, x0, x1, y0, y1);
        double best = 0.0;
        gatherMacros(own_, x0, x1, y0, y1, s);
        for (int t : own_.cand) {
            const Box b = rect(t, cx_[t], cy_[t]);
            const double o = overlapOf(a, b);
            if (o > best) {
                best = o;
                out = b;
            }
        }
        if (!blocks_.empty()) {
            gatherBlocks(own_, x0, x1, y0, y1);
            for (int b : own_.bcand) {
                const double o = overlapOf(a, blocks_[b]);
                if (o > best) {
                    best = o;
                    out = blocks_[b];
                }
            }
        }
        return best > 0.0;
    }

    // Macros (other than skip) overlapping r with positive area, ascending.
    void overlapping(const Box& r, int skip, std::vector<int>& out) const {
        int x0, x1, y0, y1;
        range(r, x0, x1, y0, y1);
        out.clear();
        const int tick = ++own_.tick;
        for (int gx = x0; gx <= x1; ++gx)
            for (int gy = y0; gy <= y1; ++gy)
                for (int t : cells_[idx(gx, gy)]) {
                    if (t == skip || own_.seen[t] == tick) continue;
                    own_.seen[t] = tick;
                    if (overlapOf(r, rect(t, cx_[t], cy_[t])) > 0.f) out.push_back(t);
                }
        std::sort(out.begin(), out.end());
    }

    // Total overlap, pairs in index order (independent of the index).
    double total() const {
        double sum = 0.0;
        for (size_t s = 0; s < w_.size(); ++s) {
            const Box a = rect(static_cast<int>(s), cx_[s], cy_[s]);
            for (size_t t = s + 1; t < w_.size(); ++t)
                sum += overlapOf(a, rect(static_cast<int>(t), cx_[t], cy_[t]));
            for (const Box& b : blocks_) sum += overlapOf(a, b);
        }
        return sum;
    }

    // Largest difference between at() and a full scan (the index self-check).
    double selfCheck() const {
        double worstDev = 0.0;
        const int n = static_cast<int>(w_.size());
        for (int s = 0; s < n; ++s) {
            const Box a = rect(s, cx_[s], cy_[s]);
            double brute = 0.0;
            for (int t = 0; t < n; ++t)
                if (t != s) brute += overlapOf(a, rect(t, cx_[t], cy_[t]));
            for (const Box& b : blocks_) brute += overlapOf(a, b);
            worstDev = std::max(worstDev, std::fabs(brute - at(s, cx_[s], cy_[s])));
        }
        return worstDev;
    }

private:
    size_t idx(int gx, int gy) const { return static_cast<size_t>(gx) * dim_ + gy; }
    void range(const Box& r, int& x0, int& x1, int& y0, int& y1) const {
        x0 = std::clamp(static_cast<int>((r.lx - ox_) / cw_), 0, dim_ - 1);
        x1 = std::clamp(static_cast<int>((r.hx - ox_) / cw_), 0, dim_ - 1);
        y0 = std::clamp(static_cast<int>((r.ly - oy_) / ch_), 0, dim_ - 1);
        y1 = std::clamp(static_cast<int>((r.hy - oy_) / ch_), 0, dim_ - 1);
    }
    void gatherMacros(Scratch& sc, int x0, int x1, int y0, int y1, int skip) const {
        sc.cand.clear();
        const int tick = ++sc.tick;
        for (int gx = x0; gx <= x1; ++gx)
            for (int gy = y0; gy <= y1; ++gy)
                for (int t : cells_[idx(gx, gy)]) {
                    if (t == skip || sc.seen[t] == tick) continue;
                    sc.seen[t] = tick;
                    sc.cand.push_back(t);
                }
        std::sort(sc.cand.begin(), sc.cand.end());
    }
    void gatherBlocks(Scratch& sc, int x0, int x1, int y0, int y1) const {
        sc.bcand.clear();
        for (int gx = x0; gx <= x1; ++gx)
            for (int gy = y0; gy <= y1; ++gy)
                for (int b : blockCells_[idx(gx, gy)]) {
                    if (sc.blockSeen[b] == sc.tick) continue;
                    sc.blockSeen[b] = sc.tick;
                    sc.bcand.push_back(b);
                }
        std::sort(sc.bcand.begin(), sc.bcand.end());
    }
    void insert(int s) {
        int x0, x1, y0, y1;
        range(rect(s, cx_[s], cy_[s]), x0, x1, y0, y1);
        for (int gx = x0; gx <= x1; ++gx)
            for (int gy = y0; gy <= y1; ++gy) cells_[idx(gx, gy)].push_back(s);
    }
    void erase(int s) {
        int x0, x1, y0, y1;
        range(rect(s, cx_[s], cy_[s]), x0, x1, y0, y1);
        for (int gx = x0; gx <= x1; ++gx)
            for (int gy = y0; gy <= y1; ++gy) {
                std::vector<int>& c = cells_[idx(gx, gy)];
                for (size_t k = 0; k < c.size(); ++k)
                    if (c[k] == s) {
                        c[k] = c.back();
                        c.pop_back();
                        break;
                    }
            }
    }

    std::vector<float> w_, h_, cx_, cy_;
    std::vector<Box> blocks_;
    int dim_ = 1;
    float ox_ = 0.f, oy_ = 0.f, cw_ = 1.f, ch_ = 1.f;
    std::vector<std::vector<int>> cells_, blockCells_;
    mutable Scratch own_;
};

// Snap a macro centre to the site / row grid, clamped into the core.
struct Grid {
    float lx, ly, hx, hy, site, row;
    void operator()(const OverlapIndex& om, int s, float& x, float& y) const {
        const float hw = 0.5f * om.width(s), hh = 0.5f * om.height(s);
        float l = x - hw, b = y - hh;
        l = lx + std::round((l - lx) / site) * site;
        b = ly + std::round((b - ly) / row) * row;
        l = std::min(std::max(l, lx), std::max(lx, hx - 2.f * hw));
        b = std::min(std::max(b, ly), std::max(ly, hy - 2.f * hh));
        x = l + hw;
        y = b + hh;
    }
};

// Positions where the macro touches a nearby rectangle (both sides of each,
// and the core edges), the start and the current point, within radius of
// the start (repair_candidates.h:14-45): the grid xs x ys, in that order,
// less the points that contactOk() rejects.
struct ContactAxes {
    std::vector<float> xs, ys;
};
ContactAxes contactAxes(float sx, float sy, float curX, float curY, float w, float h,
                        const Box& core, const std::vector<Box>& obstacles, double radius) {
    const float hw = w * .5f, hh = h * .5f;
    ContactAxes a;
    a.xs = {sx, curX, core.lx + hw, core.hx - hw};
    a.ys = {sy, curY, core.ly + hh, core.hy - hh};
    for (const Box& b : obstacles) {
        if (b.hx < sx - radius - hw || b.lx > sx + radius + hw || b.hy < sy - radius - hh ||
            b.ly > sy + radius + hh)
            continue;
        a.xs.push_back(b.lx - hw);
        a.xs.push_back(b.hx + hw);
        a.ys.push_back(b.ly - hh);
        a.ys.push_back(b.hy + hh);
    }
    for (auto* v : {&a.xs, &a.ys}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    return a;
}
inline bool contactOk(float x, float y, float sx, float sy, float w, float h, const Box& core,
                      double radius) {
    const float hw = w * .5f, hh = h * .5f;
    return !(x < core.lx + hw || x > core.hx - hw || y < core.ly + hh || y > core.hy - hh ||
             std::hypot(double(x - sx), double(y - sy)) > radius);
}
std::vector<std::pair<float, float>> contacts(float sx, float sy, float curX, float curY, float w,
                                              float h, const Box& core,
                                              const std::vector<Box>& obstacles, double radius) {
    const ContactAxes a = contactAxes(sx, sy, curX, curY, w, h, core, obstacles, radius);
    std::vector<std::pair<float, float>> out;
    for (float x : a.xs)
        for (float y : a.ys)
            if (contactOk(x, y, sx, sy, w, h, core, radius)) out.emplace_back(x, y);
    return out;
}

// Config::random_seed's generator with the reference's Rng draws
// (src/util/rng.h:12-28).
class Draw {
public:
    explicit Draw(uint64_t seed) : gen_(static_cast<std::mt19937::result_type>(seed)) {}
    float real(float lo, float hi) { return std::uniform_real_distribution<float>(lo, hi)(gen_); }
    int integer(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(gen_); }

private:
    std::mt19937 gen_;
};

// A fixed set of worker threads running f(i) for i in [0, count) with the
// calling thread joining in; run() returns when every i is done.  Used for
// the candidate evaluations of the grid repair, which are independent
// queries of the overlap index (each thread with its own scratch).
class WorkerPool {
public:
    explicit WorkerPool(int threads) {
        for (int t = 1; t < threads; ++t) workers_.emplace_back([this, t] { loop(t); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& w : workers_) w.join();
    }
    int size() const { return static_cast<int>(workers_.size()) + 1; }
    // f(i, thread) for every i; thread in [0, size()).
    void run(int count, const std::function<void(int, int)>& f) {
        if (workers_.empty() || count < 2) {
            for (int i = 0; i < count; ++i) f(i, 0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &f;
            count_ = count;
            next_.store(0);
            busy_ = static_cast<int>(workers_.size());
            ++gen_;
        }
        cv_.notify_all();
        work(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return busy_ == 0; });
        job_ = nullptr;
    }

private:
    void work(int t) {
        for (int i = next_.fetch_add(1); i < count_; i = next_.fetch_add(1)) (*job_)(i, t);
    }
    void loop(int t) {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work(t);
            std::lock_guard<std::mutex> lk(m_);
            if (--busy_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)>* job_ = nullptr;
    int count_ = 0, busy_ = 0;
    unsigned gen_ = 0;
    std::atomic<int> next_{0};
    bool stop_ = false;
};

// Threads of the grid repair: the host's hardware threads, at most 16.
inline int repairThreads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::clamp(hw == 0 ? 1u : hw, 1u, 16u));
}

class MacroLegalizer {
public:
    MacroLegalizer(View v, uint8_t* flags, uint8_t* orient, const sp_mlg_config& cfg, float site,
                   float rowH, sp_metrics* sink, const char* const* names)
        : v_(v), flags_(flags), orient_(orient), cfg_(cfg), sink_(sink), names_(names) {
        grid_ = Grid{v.d.core_lx, v.d.core_ly, v.d.core_hx, v.d.core_hy, site,
                     rowH > 0.f ? rowH : 1.f};
    }

    sp_mlg_result run();

private:
    double dispOf(int s, float px, float py) const {
        return std::hypot(static_cast<double>(px - x0_[s]), static_cast<double>(py - y0_[s]));
    }
    void number(const std::string& k, double val) {
        if (sink_) check(sp_metrics_set_number(sink_, k.c_str(), val));
    }
    void text(const std::string& k, const std::string& val) {
        if (sink_) check(sp_metrics_set_string(sink_, k.c_str(), val.c_str()));
    }
    static void check(int st) {
        if (st == SP_HIP_ENONFINITE) throw NonFinite(g_lastError);
        if (st != SP_HIP_OK) throw InvalidArg(g_lastError);
    }
    std::string nameOf(int node) const {
        return names_ ? std::string(names_[node]) : std::to_string(node);
    }
    void record(const std::string& step);
    void gridRepair(double zeroCap, int& repaired);
    void separate(bool preferSmall, int& repaired);
    void separateBoth(int& repaired);
    void chains(int& chained);

    View v_;
    uint8_t* flags_;
    uint8_t* orient_;
    sp_mlg_config cfg_;
    sp_metrics* sink_;
    const char* const* names_;
    Grid grid_{};
    std::vector<int> macros_;
    std::vector<float> x0_, y0_;
    OverlapIndex om_;
    double overlap_ = 0.0, refLen_ = 1.0;
};

void MacroLegalizer::record(const std::string& step) {
    const std::vector<float>& cx = om_.xs();
    const std::vector<float>& cy = om_.ys();
    const size_t nM = macros_.size();
    double sum = 0.0, mx = -1.0;
    size_t worst = 0;
    for (size_t s = 0; s < nM; ++s) {
        const double dd = dispOf(static_cast<int>(s), cx[s], cy[s]);
        sum += dd;
        if (dd > mx) {
            mx = dd;
            worst = s;
        }
    }
    const std::string key = "mlg_" + step + "_";
    number(key + "overlap", om_.total());
    number(key + "disp_mean", sum / nM);
    number(key + "disp_max", mx);
    text(key + "worst_macro", nameOf(macros_[worst]));
    number(key + "original_x", x0_[worst]);
    number(key + "original_y", y0_[worst]);
    number(key + "x", cx[worst]);
    number(key + "y", cy[worst]);
}

// Grid-and-contact search around each overlapping macro, the most
// overlapped first (macro_sa.cpp:835-947).
void MacroLegalizer::gridRepair(double zeroCap, int& repaired) {
    const std::vector<float>& cx = om_.xs();
    const std::vector<float>& cy = om_.ys();
    const size_t nM = macros_.size();
    const Box core{v_.d.core_lx, v_.d.core_ly, v_.d.core_hx, v_.d.core_hy};
    std::vector<size_t> order(nM);
    for (size_t s = 0; s < nM; ++s) order[s] = s;
    // The candidates of one macro are independent queries of the index: they
    // are evaluated by all threads into arrays, then scanned in the sequential
    // order, so the choice is the one a single thread makes.
    WorkerPool pool(repairThreads());
    std::vector<OverlapIndex::Scratch> scratch(pool.size());
    for (auto& sc : scratch) sc = om_.scratch();
    struct Cand {
        float x, y;
        double d, o;
    };
    std::vector<Cand> cand;
    for (int pass = 0; pass < 12 && overlap_ > 0.0; ++pass) {
        std::vector<double> ov(nM);
        for (size_t s = 0; s < nM; ++s) ov[s] = om_.at(static_cast<int>(s), cx[s], cy[s]);
        std::stable_sort(order.begin(), order.end(), [&ov](size_t a, size_t b) { return ov[a] > ov[b]; });
        for (size_t s : order) {
            const int si = static_cast<int>(s);
            const double cur = om_.at(si, cx[s], cy[s]);
            if (cur <= 0.0) continue;
            float zx = 0.f, zy = 0.f, fx = 0.f, fy = 0.f;
            double zDist = std::numeric_limits<double>::infinity(), fOv = cur;
            bool haveZero = false, haveFb = false;
            const float stepX = std::max(om_.width(si) * 0.25f, grid_.site);
            const float stepY = std::max(om_.height(si) * 0.25f, grid_.row);
            const int span = 12 + 2 * pass;
            const double side = std::max(om_.width(si), om_.height(si));
            const double zeroLimit = zeroCap * side;
            const double maxDisp = std::min(static_cast<double>(2 + pass) * side, zeroLimit);
            const int edge = 2 * span + 1;
            cand.resize(static_cast<size_t>(edge) * edge);
            const float sx = cx[s], sy = cy[s];
            pool.run(edge * edge, [&](int k, int t) {
                const int ry = k / edge - span, rx = k % edge - span;
                float nx = sx + static_cast<float>(rx) * stepX;
                float ny = sy + static_cast<float>(ry) * stepY;
                grid_(om_, si, nx, ny);
                cand[k] = Cand{nx, ny, dispOf(si, nx, ny), om_.atWith(scratch[t], si, nx, ny)};
            });
            for (const Cand& c : cand) {   // ry outer, rx inner
                if (c.o <= 0.0 && c.d <= zeroLimit) {
                    if (c.d < zDist) {
                        zDist = c.d;
                        zx = c.x;
                        zy = c.y;
                        haveZero = true;
                    }
                } else if (!haveZero && c.d <= maxDisp && c.o < fOv - 1e-9) {
                    fOv = c.o;
                    fx = c.x;
                    fy = c.y;
                    haveFb = true;
                }
            }
            std::vector<Box> obstacles = om_.blocks();
            for (size_t t = 0; t < nM; ++t)
                if (t != s) obstacles.push_back(om_.rect(static_cast<int>(t), cx[t], cy[t]));
            const double radius = std::min(zeroLimit, static_cast<double>(2 + pass) * side);
            // The sequential scan over the contact points (x outer, y inner,
            // then the 3 x 3 neighbours) keeps the first candidate of least
            // displacement below zDist with zero overlap: the minimum of
            // (displacement, position in the scan) over those candidates,
            // found here per thread over rows of xs and then combined.
            const float w = om_.width(si), h = om_.height(si), ox = x0_[s], oy = y0_[s];
            const ContactAxes ax = contactAxes(ox, oy, cx[s], cy[s], w, h, core, obstacles, radius);
            const long ny9 = static_cast<long>(ax.ys.size()) * 9;
            struct Best {
                double d;
                long at;
                float x, y;
            };
            std::vector<Best> best(pool.size(), Best{zDist, -1, 0.f, 0.f});
            pool.run(static_cast<int>(ax.xs.size()), [&](int xi, int t) {
                Best& b = best[t];
                const float px = ax.xs[xi];
                for (size_t yi = 0; yi < ax.ys.size(); ++yi) {
                    const float py = ax.ys[yi];
                    if (!contactOk(px, py, ox, oy, w, h, core, radius)) continue;
                    for (int dx = -1; dx <= 1; ++dx)
                        for (int dy = -1; dy <= 1; ++dy) {
                            const long at = xi * ny9 + static_cast<long>(yi) * 9 + (dx + 1) * 3 + dy + 1;
                            float nx = px + dx * grid_.site, ny = py + dy * grid_.row;
                            grid_(om_, si, nx, ny);
                            const double dd = dispOf(si, nx, ny);
                            if (!(dd <= zeroLimit) || !(dd < zDist)) continue;
                            if (dd > b.d || (dd == b.d && b.at >= 0 && at > b.at)) continue;
                            if (om_.atWith(scratch[t], si, nx, ny) <= 0.0) b = Best{dd, at, nx, ny};
                        }
                }
            });
            const Best* pick = nullptr;
            for (const Best& b : best)
                if (b.at >= 0 && (!pick || b.d < pick->d || (b.d == pick->d && b.at < pick->at)))
                    pick = &b;
            if (pick) {
                zDist = pick->d;
                zx = pick->x;
                zy = pick->y;
                haveZero = true;
            }
            if (haveZero) {
                om_.move(si, zx, zy);
                ++repaired;
            } else if (haveFb) {
                om_.move(si, fx, fy);
                ++repaired;
            }
        }
        overlap_ = om_.total();
    }
}

// Minimum separating shift away from each macro's worst neighbour
// (macro_sa.cpp:950-998).
void MacroLegalizer::separate(bool preferSmall, int& repaired) {
    const std::vector<float>& cx = om_.xs();
    const std::vector<float>& cy = om_.ys();
    const size_t nM = macros_.size();
    for (int pass = 0; pass < 30 && overlap_ > 0.0; ++pass) {
        bool moved = false;
        for (size_t s = 0; s < nM; ++s) {
            const int si = static_cast<int>(s);
            if (om_.at(si, cx[s], cy[s]) <= 0.0) continue;
            const Box a = om_.rect(si, cx[s], cy[s]);
            Box b{};
            if (!om_.worst(si, cx[s], cy[s], b)) continue;
            const float shift[4][2] = {{b.lx - a.hx, 0.f}, {b.hx - a.lx, 0.f},
                                       {0.f, b.ly - a.hy}, {0.f, b.hy - a.ly}};
            double bestOv = om_.at(si, cx[s], cy[s]);
            const double curOv = bestOv;
            double bestDisp = std::numeric_limits<double>::infinity();
            float bx = cx[s], by = cy[s];
            for (const auto& dv : shift) {
                float nx = cx[s] + dv[0], ny = cy[s] + dv[1];
                grid_(om_, si, nx, ny);
                const double o = om_.at(si, nx, ny);
                const double dd = dispOf(si, nx, ny);
                const bool tie = preferSmall && std::fabs(o - bestOv) <= 1e-9 && dd < bestDisp;
                if (o < curOv - 1e-9 && (o < bestOv - 1e-9 || tie)) {
                    bestOv = o;
                    bestDisp = dd;
                    bx = nx;
                    by = ny;
                }
            }
            if (bx != cx[s] || by != cy[s]) {
                om_.move(si, bx, by);
                ++repaired;
                moved = true;
            }
        }
        const double after = om_.total();
        if (!moved || after >= overlap_) {
            overlap_ = after;
            break;
        }
        overlap_ = after;
    }
}

// Both tie rules from the same start, the second kept only if it is not
// worse in overlap, total or largest displacement (macro_sa.cpp:1000-1035).
void MacroLegalizer::separateBoth(int& repaired) {
    if (overlap_ <= 0.0) return;
    const size_t nM = macros_.size();
    const std::vector<float> inX = om_.xs(), inY = om_.ys();
    const double inOverlap = overlap_;
    const int inRepaired = repaired;
    separate(false, repaired);
    const std::vector<float> oldX = om_.xs(), oldY = om_.ys();
    const double oldOverlap = overlap_;
    const int oldRepaired = repaired;
    auto displacement = [&]() {
        double sum = 0.0, mx = 0.0;
        for (size_t s = 0; s < nM; ++s) {
            const double dd = dispOf(static_cast<int>(s), om_.xs()[s], om_.ys()[s]);
            sum += dd;
            mx = std::max(mx, dd);
        }
        return std::pair<double, double>{sum, mx};
    };
    const auto oldDisp = displacement();
    for (size_t s = 0; s < nM; ++s) om_.move(static_cast<int>(s), inX[s], inY[s]);
    overlap_ = inOverlap;
    repaired = inRepaired;
    separate(true, repaired);
    const auto newDisp = displacement();
    if (overlap_ > oldOverlap || newDisp.first > oldDisp.first || newDisp.second > oldDisp.second) {
        for (size_t s = 0; s < nM; ++s) om_.move(static_cast<int>(s), oldX[s], oldY[s]);
        overlap_ = oldOverlap;
        repaired = oldRepaired;
    }
}

// Chain pushes: the pushed macro carries the macros ahead of it by the
// amounts needed, as one transaction (macro_sa.cpp:1043-1166).
void MacroLegalizer::chains(int& chained) {
    const std::vector<float>& cx = om_.xs();
    const std::vector<float>& cy = om_.ys();
    const size_t nM = macros_.size();
    const float clx = v_.d.core_lx, cly = v_.d.core_ly, chx = v_.d.core_hx, chy = v_.d.core_hy;
    for (int pass = 0; pass < 20 && overlap_ > 0.0; ++pass) {
        constexpr int kMaxChain = 64;
        std::vector<float> need(nM, 0.f);
        std::vector<int> mark(nM, -1);
        std::vector<int> chain, hits;
        std::vector<float> keepX, keepY;
        int tick = 0;
        auto build = [&](int s0, int axis, int sign, float shift0) {
            const float unit = axis == 0 ? grid_.site : grid_.row;
            auto up = [unit](float val) { return std::ceil(val / unit - 1e-4f) * unit; };
            const float cap = shift0 + 2.f * static_cast<float>(refLen_);
            ++tick;
            chain.assign(1, s0);
            mark[s0] = tick;
            need[s0] = up(shift0);
            for (size_t qi = 0; qi < chain.size(); ++qi) {
                if (qi > 4 * kMaxChain) return false;
                const int t = chain[qi];
                const float dd = need[t];
                const Box a = om_.rect(t, cx[t] + (axis == 0 ? sign * dd : 0.f),
                                       cy[t] + (axis == 1 ? sign * dd : 0.f));
                if (a.lx < clx - 1e-3f || a.hx > chx + 1e-3f || a.ly < cly - 1e-3f ||
                    a.hy > chy + 1e-3f)
                    return false;
                om_.overlapping(a, t, hits);
                const float ct = axis == 0 ? cx[t] : cy[t];
                for (int u : hits) {
                    const float cu = axis == 0 ? cx[u] : cy[u];
                    if ((cu - ct) * static_cast<float>(sign) <= 0.f) continue;
                    const Box b = om_.rect(u, cx[u], cy[u]);
                    const float req = up(sign > 0 ? (axis == 0 ? a.hx - b.lx : a.hy - b.ly)
                                                  : (axis == 0 ? b.hx - a.lx : b.hy - a.ly));
                    if (req > cap) return false;
                    const bool seen = mark[u] == tick;
                    if (seen && req <= need[u]) continue;
                    if (!seen) {
                        if (static_cast<int>(chain.size()) >= kMaxChain) return false;
                        mark[u] = tick;
                    }
                    need[u] = req;
                    chain.push_back(u);
                }
            }
            std::sort(chain.begin(), chain.end());
            chain.erase(std::unique(chain.begin(), chain.end()), chain.end());
            return true;
        };
        auto chainOverlap = [&]() {
            double sum = 0.0;
            for (int u : chain) sum += om_.at(u, cx[u], cy[u]);
            for (size_t i = 0; i < chain.size(); ++i)
                for (size_t j = i + 1; j < chain.size(); ++j) {
                    const int u = chain[i], w = chain[j];
                    sum -= overlapOf(om_.rect(u, cx[u], cy[u]), om_.rect(w, cx[w], cy[w]));
                }
            return sum;
        };
        auto apply = [&](int axis, int sign, bool keep) {
            const double before = chainOverlap();
            keepX.resize(chain.size());
            keepY.resize(chain.size());
            for (size_t k = 0; k < chain.size(); ++k) {
                const int u = chain[k];
                keepX[k] = cx[u];
                keepY[k] = cy[u];
                const float dd = static_cast<float>(sign) * need[u];
                om_.move(u, keepX[k] + (axis == 0 ? dd : 0.f), keepY[k] + (axis == 1 ? dd : 0.f));
            }
            const double gain = before - chainOverlap();
            if (!keep)
                for (size_t k = 0; k < chain.size(); ++k) om_.move(chain[k], keepX[k], keepY[k]);
            return gain;
        };
        bool moved = false;
        for (size_t s = 0; s < nM; ++s) {
            const int si = static_cast<int>(s);
            if (om_.at(si, cx[s], cy[s]) <= 0.0) continue;
            Box b{};
            if (!om_.worst(si, cx[s], cy[s], b)) continue;
            const Box a = om_.rect(si, cx[s], cy[s]);
            const float amt[4] = {a.hx - b.lx, b.hx - a.lx, a.hy - b.ly, b.hy - a.ly};
            const int axisOf[4] = {0, 0, 1, 1}, signOf[4] = {-1, 1, -1, 1};
            int bestDir = -1;
            double bestGain = 1e-6;
            for (int k = 0; k < 4; ++k) {
                if (!(amt[k] > 0.f) || !build(si, axisOf[k], signOf[k], amt[k])) continue;
                const double g = apply(axisOf[k], signOf[k], false);
                if (g > bestGain) {
                    bestGain = g;
                    bestDir = k;
                }
            }
            if (bestDir < 0) continue;
            build(si, axisOf[bestDir], signOf[bestDir], amt[bestDir]);
            apply(axisOf[bestDir], signOf[bestDir], true);
            chained += static_cast<int>(chain.size());
            moved = true;
        }
        const double after = om_.total();
        if (!moved || after >= overlap_) {
            overlap_ = after;
            break;
        }
        overlap_ = after;
    }
}

sp_mlg_result MacroLegalizer::run() {
    sp_mlg_result r{};
    const sp_hip_design& d = v_.d;
    for (int i = 0; i < d.num_movable; ++i)
        if ((flags_[i] & F_MACRO) && !(flags_[i] & F_FIXED)) macros_.push_back(i);
    r.num_macros = static_cast<int>(macros_.size());
    if (macros_.empty()) return r;
    const auto t0clock = std::chrono::steady_clock::now();
    auto msSince = [&](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };

    const size_t nM = macros_.size();
    x0_.resize(nM);
    y0_.resize(nM);
    std::vector<uint8_t> flip(nM, 0);
    for (size_t s = 0; s < nM; ++s) {
        x0_[s] = v_.x[macros_[s]];
        y0_[s] = v_.y[macros_[s]];
    }
    om_.build(v_, flags_, macros_, x0_, y0_);
    MacroWirelength iwl;
    iwl.build(v_, macros_);
    for (size_t s = 0; s < nM; ++s) {
        float px = om_.xs()[s], py = om_.ys()[s];
        grid_(om_, static_cast<int>(s), px, py);
        om_.move(static_cast<int>(s), px, py);
    }
    const std::vector<float>& cx = om_.xs();
    const std::vector<float>& cy = om_.ys();
    if (nM <= 8000) {
        const double dev = om_.selfCheck();
        if (!(std::isfinite(dev) && dev <= 1e-3))
            throw InvalidArg("mLG: spatial index disagrees with the full scan");
    }
    r.hpwl_before = hpwlOf(v_);
    r.overlap_before = om_.total();

    double macroArea = 0.0;
    for (int m : macros_) macroArea += static_cast<double>(d.node_w[m]) * static_cast<double>(d.node_h[m]);
    double wl = iwl.total(cx, cy, flip);
    refLen_ = std::sqrt(std::max(macroArea / static_cast<double>(nM), 1.0));
    const double refOverlap = std::max(macroArea, 1.0), refWl = std::max(wl, 1.0);
    const double refDisp = refLen_ * static_cast<double>(nM);
    const double kAlpha = 8.0, kBeta = 1.0, kDelta = static_cast<double>(cfg_.disp_weight);
    double overlap = r.overlap_before, disp = 0.0;
    auto cost = [&](double ov, double w, double dd) {
        return kAlpha * ov / refOverlap + kBeta * w / refWl + kDelta * dd / refDisp;
    };
    for (size_t s = 0; s < nM; ++s) disp += dispOf(static_cast<int>(s), cx[s], cy[s]);

    std::vector<int> hot;
    auto rebuildHot = [&] {
        hot.clear();
        for (size_t s = 0; s < nM; ++s)
            if (om_.at(static_cast<int>(s), cx[s], cy[s]) > 0.0) hot.push_back(static_cast<int>(s));
    };
    rebuildHot();
    const int nHot0 = static_cast<int>(hot.size());
    const int moves = cfg_.moves > 0 ? cfg_.moves : std::max(20000, nHot0 * 300);
    r.hot_macros = nHot0;
    Draw rng(cfg_.random_seed ^ 0x9E3779B97F4A7C15ull);
    const int last = static_cast<int>(nM) - 1;

    // start temperature from the median worsening of probe moves
    const double accept0 = std::clamp(static_cast<double>(cfg_.init_accept), 1e-6, 1.0 - 1e-6);
    double typical = 0.0;
    {
        std::vector<double> worse;
        worse.reserve(2000);
        std::vector<int> pn;
        std::vector<double> pv;
        const double c0 = cost(overlap, wl, disp);
        for (int k = 0; k < 2000; ++k) {
            const int s = rng.integer(0, last);
            const float ox = cx[s], oy = cy[s];
            const float span0 = static_cast<float>(refLen_);
            float nx = ox + rng.real(-span0, span0);
            float ny = oy + rng.real(-span0, span0);
            grid_(om_, s, nx, ny);
            const double dOv = om_.at(s, nx, ny) - om_.at(s, ox, oy);
            const double dD = dispOf(s, nx, ny) - dispOf(s, ox, oy);
            om_.move(s, nx, ny);
            const double dW = iwl.delta(s, cx, cy, flip, pn, pv);
            om_.move(s, ox, oy);
            const double dc = cost(overlap + dOv, wl + dW, disp + dD) - c0;
            if (dc > 0.0) worse.push_back(dc);
        }
        if (!worse.empty()) {
            std::nth_element(worse.begin(), worse.begin() + static_cast<long>(worse.size() / 2),
                             worse.end());
            typical = worse[worse.size() / 2];
        }
    }
    const double t0 = typical > 0.0 ? -typical / std::log(accept0) : -1.0 / std::log(accept0);
    const double ratio = std::max(1e-12f, cfg_.temp_ratio);
    r.t0 = t0;
    r.typical_worsen = typical;

    std::vector<float> bestX = cx, bestY = cy;
    std::vector<uint8_t> bestFlip = flip;
    double bestCost = cost(overlap, wl, disp);
    std::vector<int> dirty, dirty2;
    std::vector<double> vals, vals2;
    int lastImprove = 0, done = 0;
    for (int it = 0; it < moves; ++it) {
        done = it + 1;
        const double progress = static_cast<double>(it) / static_cast<double>(moves);
        const double temp = t0 * std::pow(ratio, progress);
        if (it % 500 == 0 && it > 0) rebuildHot();
        if (overlap <= 0.0 && it - lastImprove > 20000) break;
        const float span = static_cast<float>(refLen_ * (0.5 * (1.0 - progress) + 0.02));
        const bool targeted = !hot.empty() && rng.real(0.f, 1.f) < 0.7f;
        const int s = targeted ? hot[rng.integer(0, static_cast<int>(hot.size()) - 1)]
                               : rng.integer(0, last);
        const int op = targeted ? 0 : rng.integer(0, 9);   // 0-6 shift, 7-8 swap, 9 flip
        const float oldX = cx[s], oldY = cy[s];
        const uint8_t oldFlip = flip[s];
        int s2 = -1;
        float oldX2 = 0.f, oldY2 = 0.f;
        double dOv = 0.0, dD = 0.0;
        if (op <= 6) {
            float nx = oldX, ny = oldY;
            Box b{};
            if (targeted && om_.worst(s, oldX, oldY, b)) {
                const Box a = om_.rect(s, oldX, oldY);
                const float cand[4][2] = {{b.lx - a.hx, 0.f}, {b.hx - a.lx, 0.f},
                                          {0.f, b.ly - a.hy}, {0.f, b.hy - a.ly}};
                int pick = 0;
                if (rng.real(0.f, 1.f) < 0.7f) {
                    float bestM = std::numeric_limits<float>::infinity();
                    for (int k = 0; k < 4; ++k) {
                        const float m = std::fabs(cand[k][0]) + std::fabs(cand[k][1]);
                        if (m < bestM) {
                            bestM = m;
                            pick = k;
                        }
                    }
                } else {
                    pick = rng.integer(0, 3);
                }
                nx = oldX + cand[pick][0];
                ny = oldY + cand[pick][1];
            } else {
                nx = oldX + rng.real(-span, span);
                ny = oldY + rng.real(-span, span);
            }
            grid_(om_, s, nx, ny);
            dOv = om_.at(s, nx, ny) - om_.at(s, oldX, oldY);
            dD = dispOf(s, nx, ny) - dispOf(s, oldX, oldY);
            om_.move(s, nx, ny);
        } else if (op <= 8 && nM >= 2) {
            do {
                s2 = rng.integer(0, last);
            } while (s2 == s);
            oldX2 = cx[s2];
            oldY2 = cy[s2];
            const double before = om_.at(s, oldX, oldY) + om_.at(s2, oldX2, oldY2) -
                                  overlapOf(om_.rect(s, oldX, oldY), om_.rect(s2, oldX2, oldY2));
            float ax = oldX2, ay = oldY2, bx = oldX, by = oldY;
            grid_(om_, s, ax, ay);
            grid_(om_, s2, bx, by);
            om_.move(s, ax, ay);
            om_.move(s2, bx, by);
            const double after = om_.at(s, ax, ay) + om_.at(s2, bx, by) -
                                 overlapOf(om_.rect(s, ax, ay), om_.rect(s2, bx, by));
            dOv = after - before;
            dD = dispOf(s, ax, ay) - dispOf(s, oldX, oldY) + dispOf(s2, bx, by) -
                 dispOf(s2, oldX2, oldY2);
        } else {
            flip[s] = static_cast<uint8_t>(1 - oldFlip);   // pins only: overlap unchanged
        }
        double dW = iwl.delta(s, cx, cy, flip, dirty, vals);
        dirty2.clear();
        vals2.clear();
        if (s2 >= 0) dW += iwl.delta(s2, cx, cy, flip, dirty2, vals2);
        const double dCost = cost(overlap + dOv, wl + dW, disp + dD) - cost(overlap, wl, disp);
        const bool accept = dCost <= 0.0 || (temp > 0.0 && rng.real(0.f, 1.f) < std::exp(-dCost / temp));
        if (accept) {
            overlap += dOv;
            wl += dW;
            disp += dD;
            iwl.commit(dirty, vals);
            if (s2 >= 0) iwl.commit(dirty2, vals2);
            ++r.accepted;
            const double c = cost(overlap, wl, disp);
            if (c < bestCost) {
                bestCost = c;
                bestX = cx;
                bestY = cy;
                bestFlip = flip;
                lastImprove = it;
            }
        } else {
            om_.move(s, oldX, oldY);
            flip[s] = oldFlip;
            if (s2 >= 0) om_.move(s2, oldX2, oldY2);
        }
    }
    r.moves = done;
    for (size_t s = 0; s < nM; ++s) om_.move(static_cast<int>(s), bestX[s], bestY[s]);
    flip = bestFlip;
    wl = iwl.total(cx, cy, flip);
    overlap_ = om_.total();
    r.overlap_after_sa = overlap_;
    r.sa_ms = msSince(t0clock);
    record("sa");

    int repaired = 0, chained = 0;
    constexpr double kLocalCap = 3.0;
    separateBoth(repaired);
    record("separate1");
    chains(chained);
    record("chain1");
    gridRepair(kLocalCap, repaired);
    record("local_grid");
    separateBoth(repaired);
    record("separate2");
    chains(chained);
    record("chain2");
    if (overlap_ > 0.0) {
        gridRepair(std::numeric_limits<double>::infinity(), repaired);
        record("unlimited_grid");
        separateBoth(repaired);
        chains(chained);
    }
    // pull the most displaced macro toward its start when legal and not
    // longer in wirelength (macro_sa.cpp:1182-1219)
    if (overlap_ <= 0.0) {
        size_t far = 0;
        double maxD = 0.0;
        for (size_t s = 0; s < nM; ++s) {
            const double dd = dispOf(static_cast<int>(s), cx[s], cy[s]);
            if (dd > maxD) {
                maxD = dd;
                far = s;
            }
        }
        const int si = static_cast<int>(far);
        std::vector<Box> obstacles = om_.blocks();
        for (size_t s = 0; s < nM; ++s)
            if (s != far) obstacles.push_back(om_.rect(static_cast<int>(s), cx[s], cy[s]));
        const Box core{d.core_lx, d.core_ly, d.core_hx, d.core_hy};
        float bx = cx[far], by = cy[far];
        double bestD = maxD;
        for (auto [px, py] : contacts(x0_[far], y0_[far], cx[far], cy[far], om_.width(si),
                                      om_.height(si), core, obstacles, maxD))
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    float nx = px + dx * grid_.site, ny = py + dy * grid_.row;
                    grid_(om_, si, nx, ny);
                    const double dd = dispOf(si, nx, ny);
                    if (dd < bestD && om_.at(si, nx, ny) <= 0.0) {
                        bestD = dd;
                        bx = nx;
                        by = ny;
                    }
                }
        if (bestD < maxD) {
            iwl.total(cx, cy, flip);
            const float ox = cx[far], oy = cy[far];
            om_.move(si, bx, by);
            const double delta = iwl.delta(si, cx, cy, flip, dirty, vals);
            if (delta > 0.0) om_.move(si, ox, oy);
            number("mlg_postlegal_accepted", delta <= 0.0 ? 1 : 0);
        }
    }
    record("postlegal");
    record("final");
    r.repaired = repaired;
    r.chained = chained;

    // results back: positions, mirrored pins and orientations, macros fixed
    for (size_t s = 0; s < nM; ++s) {
        const int m = macros_[s];
        v_.x[m] = cx[s];
        v_.y[m] = cy[s];
        if (flip[s]) {
            for (int k = d.node2pin_start[m]; k < d.node2pin_start[m + 1]; ++k) {
                const int p = d.flat_node2pin[k];
                v_.offx[p] = -v_.offx[p];
            }
            if (orient_) {
                uint8_t& o = orient_[m];
                if (o == 0) o = 4;
                else if (o == 4) o = 0;
                else if (o == 2) o = 6;
                else if (o == 6) o = 2;
            }
        }
        flags_[m] |= F_FIXED;
    }
    r.overlap_after = om_.total();
    r.hpwl_after = hpwlOf(v_);
    if (sink_) check(sp_metrics_record_duration(sink_, "mLG", msSince(t0clock)));
    return r;
}

}  // namespace
}  // namespace sphip

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_macro_legalize(const sp_hip_design* d, const float* net_weight, float site_width,
                      float row_height, const sp_mlg_config* cfg, float* x, float* y,
                      uint8_t* node_flags, float* pin_offset_x, uint8_t* orient,
                      sp_metrics* sink, const char* const* node_names, sp_mlg_result* result) {
    return guarded([&] {
        if (!d || !cfg || !x || !y || !node_flags || !pin_offset_x || !result)
            throw InvalidArg("null argument");
        if (!(site_width > 0.f)) throw InvalidArg("site width must be positive");
        if (d->num_pins > 0 && (!d->pin2node || !d->pin2net || !d->pin_offset_y ||
                                !d->flat_net2pin || !d->flat_node2pin))
            throw InvalidArg("null pin array");
        if (!d->net2pin_start || !d->node2pin_start) throw InvalidArg("null CSR array");
        sphip::View v{*d, net_weight, x, y, pin_offset_x};
        sphip::MacroLegalizer lg(v, node_flags, orient, *cfg, site_width, row_height, sink,
                                 node_names);
        *result = lg.run();
    });
}

}  // extern "C"
