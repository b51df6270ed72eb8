// Macro legalization before Abacus: simulated annealing over the movable
// macros, then greedy repairs (the behaviour of legalizeMacros,
// src/lg/macro_sa.cpp:493-1272, with the contact points of
// src/lg/repair_candidates.h:14-45).  Host code: one sequential Markov chain
// and repair passes whose every step depends on the one before; only the
// candidate evaluations of the grid repair run on several host threads.
//
// The output must equal the reference's bit for bit on the same design and
// seed.  What fixes the bits is kept: the random draws and their order, the
// cost terms, every float/double expression, the order of every
// floating-point sum, the scan order and tie rules of every choice.  The data
// structures are this file's own:
//  * Placement holds the macro geometry and positions.  Its spatial index
//    files each macro under one cell (its centre) of a grid whose cell side
//    fits the macro (grids of doubling cell size, one per size tier), and
//    the fixed blockages in a static cell list.  Every query returns the
//    exact set of positively overlapping objects, which is then sorted by
//    index: a sum over that set in index order is the reference's sum, whose
//    zero terms change nothing.
//  * NetModel keeps the wirelength of the nets that touch macros: per net
//    the box of its other pins and the pin-offset box of each macro on it.
//  * Legalizer runs the phases: set-up and start temperature, the annealing
//    chain, the repair ladder, the pull-back of the farthest macro and the
//    write-back.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "abi.h"

#pragma STDC FP_CONTRACT OFF

namespace sphip {
namespace {

using Clock = std::chrono::steady_clock;
constexpr float kFloatInf = std::numeric_limits<float>::infinity();
constexpr double kDoubleInf = std::numeric_limits<double>::infinity();

double msSince(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

struct Rect4 {
    float lx, ly, hx, hy;
};

// Overlap area of two rectangles, 0 unless both extents are positive
// (overlapArea, src/db/geometry.h:35-39).
inline float cover(const Rect4& a, const Rect4& b) {
    const float w = std::min(a.hx, b.hx) - std::max(a.lx, b.lx);
    const float h = std::min(a.hy, b.hy) - std::max(a.ly, b.ly);
    return (w > 0.f && h > 0.f) ? w * h : 0.f;
}

// The reference's random source (src/util/rng.h:12-29): std::mt19937 seeded
// with the low 32 bits of the seed; each draw uses a fresh libstdc++
// distribution (float on [lo, hi), int on [lo, hi]).
class Draws {
public:
    explicit Draws(uint64_t seed) : gen_(static_cast<std::mt19937::result_type>(seed)) {}
    float real(float lo, float hi) { return std::uniform_real_distribution<float>(lo, hi)(gen_); }
    int pick(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(gen_); }

private:
    std::mt19937 gen_;
};

// Site / row alignment of a macro's lower-left corner, clamped into the core
// (the reference snaps every proposed position, macro_sa.cpp:470-485).
struct SiteSnap {
    float lx, ly, hx, hy, site, row;
    void apply(float hw, float hh, float& x, float& y) const {
        float l = x - hw, b = y - hh;
        l = lx + std::round((l - lx) / site) * site;
        b = ly + std::round((b - ly) / row) * row;
        l = std::min(std::max(l, lx), std::max(lx, hx - 2.f * hw));
        b = std::min(std::max(b, ly), std::max(ly, hy - 2.f * hh));
        x = l + hw;
        y = b + hh;
    }
};

// ------------------------------------------------------------------ threads
// A fixed set of helper threads for data-parallel loops: run(count, f) calls
// f(i, lane) for every i < count, lane < lanes(), the caller being lane 0.
class Lanes {
public:
    explicit Lanes(int lanes) {
        for (int l = 1; l < lanes; ++l) threads_.emplace_back([this, l] { serve(l); });
    }
    ~Lanes() {
        {
            std::lock_guard<std::mutex> g(mu_);
            quit_ = true;
            ++round_;
        }
        wake_.notify_all();
        for (auto& t : threads_) t.join();
    }
    int lanes() const { return static_cast<int>(threads_.size()) + 1; }

    void run(int count, const std::function<void(int, int)>& f) {
        if (threads_.empty() || count < 2) {
            for (int i = 0; i < count; ++i) f(i, 0);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            task_ = &f;
            total_ = count;
            cursor_.store(0);
            pending_ = static_cast<int>(threads_.size());
            ++round_;
        }
        wake_.notify_all();
        drain(0);
        std::unique_lock<std::mutex> g(mu_);
        idle_.wait(g, [this] { return pending_ == 0; });
        task_ = nullptr;
    }

private:
    void drain(int lane) {
        for (;;) {
            const int i = cursor_.fetch_add(1);
            if (i >= total_) return;
            (*task_)(i, lane);
        }
    }
    void serve(int lane) {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                wake_.wait(g, [&] { return round_ != seen; });
                seen = round_;
                if (quit_) return;
            }
            drain(lane);
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) idle_.notify_one();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable wake_, idle_;
    const std::function<void(int, int)>* task_ = nullptr;
    std::atomic<int> cursor_{0};
    int total_ = 0, pending_ = 0;
    unsigned round_ = 0;
    bool quit_ = false;
};

// ---------------------------------------------------------------- placement
// Per-query scratch (one per thread).
struct Probe {
    std::vector<int> hits, blocks;
};

// Macro rectangles, their positions, and the fixed blockages, with the
// spatial index behind the overlap queries.  place() is the only writer of a
// position, so the index always agrees with the coordinates.
class Placement {
public:
    void build(const std::vector<float>& w, const std::vector<float>& h,
               const std::vector<float>& x, const std::vector<float>& y,
               std::vector<Rect4> blockages, const Rect4& core) {
        const int n = static_cast<int>(w.size());
        w_ = w;
        h_ = h;
        hw_.resize(n);
        hh_.resize(n);
        for (int s = 0; s < n; ++s) {
            hw_[s] = 0.5f * w_[s];
            hh_[s] = 0.5f * h_[s];
        }
        x_ = x;
        y_ = y;
        blk_ = std::move(blockages);
        ox_ = core.lx;
        oy_ = core.ly;
        const double spanX = std::max(static_cast<double>(core.hx) - core.lx, 1e-3);
        const double spanY = std::max(static_cast<double>(core.hy) - core.ly, 1e-3);
        const double span = std::max(spanX, spanY);
        // tier 0: about sqrt(2 n) cells along the longer core side; each
        // further tier doubles the cell, up to the largest macro side
        const double target = std::sqrt(2.0 * std::max(n, 1));
        const double base = std::max(span / std::clamp(target, 4.0, 256.0), 1e-3);
        double biggest = 0.0;
        for (int s = 0; s < n; ++s)
            biggest = std::max(biggest, static_cast<double>(std::max(w_[s], h_[s])));
        tiers_.clear();
        for (double cell = base;; cell *= 2.0) {
            Tier t;
            t.cell = cell;
            t.nx = static_cast<int>(std::clamp(std::ceil(spanX / cell), 1.0, 4096.0));
            t.ny = static_cast<int>(std::clamp(std::ceil(spanY / cell), 1.0, 4096.0));
            t.bins.assign(static_cast<size_t>(t.nx) * t.ny, {});
            tiers_.push_back(std::move(t));
            if (cell >= biggest || tiers_.size() >= 40) break;
        }
        tierOf_.assign(n, 0);
        binOf_.assign(n, 0);
        slotInBin_.assign(n, 0);
        for (int s = 0; s < n; ++s) {
            const double side = std::max(w_[s], h_[s]);
            int k = 0;
            while (k + 1 < static_cast<int>(tiers_.size()) && !(side <= tiers_[k].cell)) ++k;
            tierOf_[s] = k;
            Tier& t = tiers_[k];
            t.half = std::max(t.half, static_cast<double>(std::max(hw_[s], hh_[s])));
            file(s);
        }
        // blockages: a static cell list, each blockage under every cell its
        // rectangle covers
        bCell_ = std::max(span / std::clamp(std::sqrt(static_cast<double>(blk_.size())), 4.0, 256.0),
                          1e-3);
        bnx_ = static_cast<int>(std::clamp(std::ceil(spanX / bCell_), 1.0, 4096.0));
        bny_ = static_cast<int>(std::clamp(std::ceil(spanY / bCell_), 1.0, 4096.0));
        std::vector<int> count(static_cast<size_t>(bnx_) * bny_ + 1, 0);
        auto each = [&](int b, auto&& fn) {
            const Rect4& r = blk_[b];
            const int x0 = cellOf(r.lx, ox_, bCell_, bnx_), x1 = cellOf(r.hx, ox_, bCell_, bnx_);
            const int y0 = cellOf(r.ly, oy_, bCell_, bny_), y1 = cellOf(r.hy, oy_, bCell_, bny_);
            for (int i = x0; i <= x1; ++i)
                for (int j = y0; j <= y1; ++j) fn(i * bny_ + j);
        };
        for (int b = 0; b < static_cast<int>(blk_.size()); ++b)
            each(b, [&](int c) { ++count[c + 1]; });
        for (size_t c = 1; c < count.size(); ++c) count[c] += count[c - 1];
        bStart_ = count;
        bList_.assign(static_cast<size_t>(count.back()), 0);
        for (int b = 0; b < static_cast<int>(blk_.size()); ++b)
            each(b, [&](int c) { bList_[count[c]++] = b; });
    }

    int size() const { return static_cast<int>(w_.size()); }
    float x(int s) const { return x_[s]; }
    float y(int s) const { return y_[s]; }
    const std::vector<float>& xs() const { return x_; }
    const std::vector<float>& ys() const { return y_; }
    float width(int s) const { return w_[s]; }
    float height(int s) const { return h_[s]; }
    float halfW(int s) const { return hw_[s]; }
    float halfH(int s) const { return hh_[s]; }
    const std::vector<Rect4>& blockages() const { return blk_; }

    Rect4 box(int s, float px, float py) const {
        return {px - hw_[s], py - hh_[s], px + hw_[s], py + hh_[s]};
    }
    Rect4 box(int s) const { return box(s, x_[s], y_[s]); }

    void place(int s, float px, float py) {
        unfile(s);
        x_[s] = px;
        y_[s] = py;
        file(s);
    }

    // Macros other than `skip` whose rectangles overlap r with positive area
    // (cover(r, macro) > 0), ascending.
    void macrosHitting(const Rect4& r, int skip, std::vector<int>& out) const {
        out.clear();
        const double margin = 1e-4 * (std::fabs(static_cast<double>(r.lx)) +
                                      std::fabs(static_cast<double>(r.hx)) +
                                      std::fabs(static_cast<double>(r.ly)) +
                                      std::fabs(static_cast<double>(r.hy))) + 1e-3;
        for (const Tier& t : tiers_) {
            if (t.members == 0) continue;
            const double reach = t.half + margin;   // a member's centre lies within this
            const int x0 = cellOf(r.lx - reach, ox_, t.cell, t.nx);
            const int x1 = cellOf(r.hx + reach, ox_, t.cell, t.nx);
            const int y0 = cellOf(r.ly - reach, oy_, t.cell, t.ny);
            const int y1 = cellOf(r.hy + reach, oy_, t.cell, t.ny);
            for (int i = x0; i <= x1; ++i)
                for (int j = y0; j <= y1; ++j)
                    for (int u : t.bins[static_cast<size_t>(i) * t.ny + j])
                        if (u != skip && cover(r, box(u)) > 0.f) out.push_back(u);
        }
        std::sort(out.begin(), out.end());
    }
    // Blockages overlapping r with positive area, ascending.
    void blockagesHitting(const Rect4& r, std::vector<int>& out) const {
        out.clear();
        if (blk_.empty()) return;
        const int x0 = cellOf(r.lx, ox_, bCell_, bnx_), x1 = cellOf(r.hx, ox_, bCell_, bnx_);
        const int y0 = cellOf(r.ly, oy_, bCell_, bny_), y1 = cellOf(r.hy, oy_, bCell_, bny_);
        for (int i = x0; i <= x1; ++i)
            for (int j = y0; j <= y1; ++j) {
                const int c = i * bny_ + j;
                for (int k = bStart_[c]; k < bStart_[c + 1]; ++k)
                    if (cover(r, blk_[bList_[k]]) > 0.f) out.push_back(bList_[k]);
            }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }

    // Overlap of macro s placed at (px, py) with the other macros, then with
    // the blockages, each summed in index order (OverlapModel::at,
    // macro_sa.cpp:284-320).
    double overlapAt(int s, float px, float py, Probe& pr) const {
        const Rect4 a = box(s, px, py);
        double sum = 0.0;
        macrosHitting(a, s, pr.hits);
        for (int t : pr.hits) sum += cover(a, box(t));
        blockagesHitting(a, pr.blocks);
        for (int b : pr.blocks) sum += cover(a, blk_[b]);
        return sum;
    }
    double overlapAt(int s, Probe& pr) const { return overlapAt(s, x_[s], y_[s], pr); }

    // The rectangle of the object overlapping macro s at (px, py) the most:
    // macros before blockages, each in index order, a later one only when
    // strictly larger (worstRect, macro_sa.cpp:322-366).
    bool heaviest(int s, float px, float py, Rect4& out, Probe& pr) const {
        const Rect4 a = box(s, px, py);
        double best = 0.0;
        macrosHitting(a, s, pr.hits);
        for (int t : pr.hits) {
            const Rect4 b = box(t);
            const double o = cover(a, b);
            if (o > best) {
                best = o;
                out = b;
            }
        }
        blockagesHitting(a, pr.blocks);
        for (int k : pr.blocks) {
            const double o = cover(a, blk_[k]);
            if (o > best) {
                best = o;
                out = blk_[k];
            }
        }
        return best > 0.0;
    }

    // Total overlap: every macro pair once (s < t) and every macro with every
    // blockage, s-major (OverlapModel::total, macro_sa.cpp:384-393).
    double total(Probe& pr) const {
        double sum = 0.0;
        for (int s = 0; s < size(); ++s) {
            const Rect4 a = box(s);
            macrosHitting(a, s, pr.hits);
            for (int t : pr.hits)
                if (t > s) sum += cover(a, box(t));
            blockagesHitting(a, pr.blocks);
            for (int b : pr.blocks) sum += cover(a, blk_[b]);
        }
        return sum;
    }

    // The same sum as overlapAt without the index: every other macro and
    // every blockage (the self-check's independent scan).
    double scanAt(int s) const {
        const Rect4 a = box(s);
        double sum = 0.0;
        for (int t = 0; t < size(); ++t)
            if (t != s) sum += cover(a, box(t));
        for (const Rect4& b : blk_) sum += cover(a, b);
        return sum;
    }

private:
    struct Tier {
        double cell = 1.0;
        double half = 0.0;   // the largest half side of a member
        int nx = 1, ny = 1, members = 0;
        std::vector<std::vector<int>> bins;
    };
    // Cell of coordinate v on an axis from origin o with cells of side c,
    // clamped to [0, n - 1] (non-finite values to cell 0).
    static int cellOf(double v, double o, double c, int n) {
        const double t = (v - o) / c;
        if (!(t >= 0.0)) return 0;
        if (t >= n) return n - 1;
        return static_cast<int>(t);
    }
    void file(int s) {
        Tier& t = tiers_[tierOf_[s]];
        const int c = cellOf(x_[s], ox_, t.cell, t.nx) * t.ny + cellOf(y_[s], oy_, t.cell, t.ny);
        std::vector<int>& bin = t.bins[c];
        binOf_[s] = c;
        slotInBin_[s] = static_cast<int>(bin.size());
        bin.push_back(s);
        ++t.members;
    }
    void unfile(int s) {
        Tier& t = tiers_[tierOf_[s]];
        std::vector<int>& bin = t.bins[binOf_[s]];
        const int at = slotInBin_[s];
        const int last = bin.back();
        bin[at] = last;
        slotInBin_[last] = at;
        bin.pop_back();
        --t.members;
    }

    std::vector<float> w_, h_, hw_, hh_, x_, y_;
    std::vector<Rect4> blk_;
    double ox_ = 0, oy_ = 0;
    std::vector<Tier> tiers_;
    std::vector<int> tierOf_, binOf_, slotInBin_;
    double bCell_ = 1.0;
    int bnx_ = 1, bny_ = 1;
    std::vector<int> bStart_, bList_;
};

// ------------------------------------------------------------------- nets
// Wirelength of the nets with macro pins (the nets without them do not move
// during macro legalization).  A net's box is the box of its other pins
// (fixed here) joined with each macro's pin-offset box on the net shifted by
// the macro's centre, mirrored in x when the macro is flipped
// (IncrementalWirelength, macro_sa.cpp:27-197).
class NetModel {
public:
    void build(const sp_hip_design& d, const float* netW, const float* offx,
               const std::vector<int>& macros, const float* nodeX, const float* nodeY) {
        const int nets = d.num_nets;
        std::vector<int> slotOf(static_cast<size_t>(d.num_nodes) + d.num_fillers, -1);
        for (size_t s = 0; s < macros.size(); ++s) slotOf[macros[s]] = static_cast<int>(s);
        auto macroPin = [&](int p) { return slotOf[d.pin2node[p]] >= 0; };
        // the box of the pins of other nodes, and whether a macro is on the net
        other_.assign(nets, Rect4{kFloatInf, kFloatInf, -kFloatInf, -kFloatInf});
        active_.assign(nets, 0);
        for (int n = 0; n < nets; ++n)
            for (int k = d.net2pin_start[n]; k < d.net2pin_start[n + 1]; ++k) {
                const int p = d.flat_net2pin[k];
                if (macroPin(p)) {
                    active_[n] = 1;
                    continue;
                }
                const float px = nodeX[d.pin2node[p]] + offx[p];
                const float py = nodeY[d.pin2node[p]] + d.pin_offset_y[p];
                Rect4& b = other_[n];
                b.lx = std::min(b.lx, px);
                b.hx = std::max(b.hx, px);
                b.ly = std::min(b.ly, py);
                b.hy = std::max(b.hy, py);
            }
        // each macro's nets in the order of its first pin on them, with its
        // pin-offset box on each
        std::vector<std::pair<int, Rect4>> own;   // (net, box) of the current macro
        std::vector<int> where(nets, -1);
        std::vector<std::vector<std::pair<int, Rect4>>> perMacro(macros.size());
        for (size_t s = 0; s < macros.size(); ++s) {
            const int m = macros[s];
            own.clear();
            for (int k = d.node2pin_start[m]; k < d.node2pin_start[m + 1]; ++k) {
                const int p = d.flat_node2pin[k];
                const int n = d.pin2net[p];
                const float ox = offx[p], oy = d.pin_offset_y[p];
                if (where[n] < 0) {
                    where[n] = static_cast<int>(own.size());
                    own.push_back({n, Rect4{ox, oy, ox, oy}});
                } else {
                    Rect4& b = own[where[n]].second;
                    b.lx = std::min(b.lx, ox);
                    b.hx = std::max(b.hx, ox);
                    b.ly = std::min(b.ly, oy);
                    b.hy = std::max(b.hy, oy);
                }
            }
            for (const auto& e : own) where[e.first] = -1;
            perMacro[s] = own;
        }
        // per macro: its nets; per net: (macro, box) entries by macro index
        macroNetStart_.assign(macros.size() + 1, 0);
        macroNet_.clear();
        std::vector<int> perNet(static_cast<size_t>(nets) + 1, 0);
        for (size_t s = 0; s < macros.size(); ++s) {
            for (const auto& e : perMacro[s]) {
                macroNet_.push_back(e.first);
                ++perNet[e.first + 1];
            }
            macroNetStart_[s + 1] = static_cast<int>(macroNet_.size());
        }
        for (int n = 0; n < nets; ++n) perNet[n + 1] += perNet[n];
        netStart_ = perNet;
        entryMacro_.assign(static_cast<size_t>(perNet.back()), 0);
        entryBox_.assign(static_cast<size_t>(perNet.back()), Rect4{});
        for (size_t s = 0; s < macros.size(); ++s)
            for (const auto& e : perMacro[s]) {
                const int at = perNet[e.first]++;
                entryMacro_[at] = static_cast<int>(s);
                entryBox_[at] = e.second;
            }
        weight_.assign(nets, 1.f);
        if (netW)
            for (int n = 0; n < nets; ++n) weight_[n] = netW[n];
        value_.assign(nets, 0.0);
    }

    // Every active net's length at these positions, refreshing the cache;
    // summed in net order.
    double total(const std::vector<float>& x, const std::vector<float>& y,
                 const std::vector<uint8_t>& flip) {
        double sum = 0.0;
        for (size_t n = 0; n < active_.size(); ++n) {
            if (!active_[n]) continue;
            value_[n] = length(static_cast<int>(n), x, y, flip);
            sum += value_[n];
        }
        return sum;
    }

    // The change of the nets of macro s against the cache, in the macro's
    // net order; their new lengths go to `fresh` (the cache is untouched).
    double change(int s, const std::vector<float>& x, const std::vector<float>& y,
                  const std::vector<uint8_t>& flip, std::vector<std::pair<int, double>>& fresh) const {
        fresh.clear();
        double sum = 0.0;
        for (int k = macroNetStart_[s]; k < macroNetStart_[s + 1]; ++k) {
            const int n = macroNet_[k];
            const double v = length(n, x, y, flip);
            fresh.push_back({n, v});
            sum += v - value_[n];
        }
        return sum;
    }
    void accept(const std::vector<std::pair<int, double>>& fresh) {
        for (const auto& e : fresh) value_[e.first] = e.second;
    }

private:
    double length(int n, const std::vector<float>& x, const std::vector<float>& y,
                  const std::vector<uint8_t>& flip) const {
        Rect4 b = other_[n];
        for (int k = netStart_[n]; k < netStart_[n + 1]; ++k) {
            const int s = entryMacro_[k];
            const Rect4& e = entryBox_[k];
            const bool f = flip[s] != 0;
            const float lo = f ? -e.hx : e.lx, hi = f ? -e.lx : e.hx;
            b.lx = std::min(b.lx, x[s] + lo);
            b.hx = std::max(b.hx, x[s] + hi);
            b.ly = std::min(b.ly, y[s] + e.ly);
            b.hy = std::max(b.hy, y[s] + e.hy);
        }
        if (!(b.hx >= b.lx) || !(b.hy >= b.ly)) return 0.0;
        return static_cast<double>(weight_[n]) *
               (static_cast<double>(b.hx - b.lx) + static_cast<double>(b.hy - b.ly));
    }

    std::vector<Rect4> other_;
    std::vector<uint8_t> active_;
    std::vector<float> weight_;
    std::vector<double> value_;
    std::vector<int> macroNetStart_, macroNet_, netStart_, entryMacro_;
    std::vector<Rect4> entryBox_;
};

// HPWL of the whole design, nets with fewer than two pins skipped, summed in
// 64 fixed chunks added in order (computeHPWL, src/db/hpwl.cpp:13-39, with
// deterministicSum, src/numeric/reduction.h:43-71).
double designHpwl(const sp_hip_design& d, const float* x, const float* y, const float* offx) {
    const int64_t nets = d.num_nets;
    if (nets <= 0) return 0.0;
    double chunk[64];
    for (int c = 0; c < 64; ++c) {
        double s = 0.0;
        for (int64_t n = nets * c / 64; n < nets * (c + 1) / 64; ++n) {
            const int first = d.net2pin_start[n], end = d.net2pin_start[n + 1];
            if (end - first < 2) continue;
            float lx = std::numeric_limits<float>::max(), hx = -lx, ly = lx, hy = -lx;
            for (int k = first; k < end; ++k) {
                const int p = d.flat_net2pin[k];
                const float px = x[d.pin2node[p]] + offx[p];
                const float py = y[d.pin2node[p]] + d.pin_offset_y[p];
                lx = std::min(lx, px);
                hx = std::max(hx, px);
                ly = std::min(ly, py);
                hy = std::max(hy, py);
            }
            s += static_cast<double>(hx - lx) + static_cast<double>(hy - ly);
        }
        chunk[c] = s;
    }
    double total = 0.0;
    for (double c : chunk) total += c;
    return total;
}

// Contact points of a macro of w x h (repair_candidates.h:14-45): x from the
// start and current positions, the core edges and the near sides of every
// obstacle within `radius` (plus the half size) of the start, y likewise;
// each list sorted and deduplicated; the crossing points that keep the macro
// in the core and lie within `radius` of the start, x-major.  `near` holds
// the obstacles to consider (any superset of those within reach, in any
// order: the lists are sorted).
std::vector<std::pair<float, float>> contactPoints(float startX, float startY, float curX,
                                                   float curY, float w, float h,
                                                   const Rect4& core,
                                                   const std::vector<Rect4>& near,
                                                   double radius) {
    const float hw = w * .5f, hh = h * .5f;
    std::vector<float> xs{startX, curX, core.lx + hw, core.hx - hw};
    std::vector<float> ys{startY, curY, core.ly + hh, core.hy - hh};
    for (const Rect4& b : near) {
        const bool far = b.hx < startX - radius - hw || b.lx > startX + radius + hw ||
                         b.hy < startY - radius - hh || b.ly > startY + radius + hh;
        if (far) continue;
        xs.push_back(b.lx - hw);
        xs.push_back(b.hx + hw);
        ys.push_back(b.ly - hh);
        ys.push_back(b.hy + hh);
    }
    for (auto* v : {&xs, &ys}) {
        std::sort(v->begin(), v->end());
        v->erase(std::unique(v->begin(), v->end()), v->end());
    }
    std::vector<std::pair<float, float>> pts;
    for (float px : xs)
        for (float py : ys) {
            const bool inCore = !(px < core.lx + hw || px > core.hx - hw || py < core.ly + hh ||
                                  py > core.hy - hh);
            if (inCore &&
                !(std::hypot(static_cast<double>(px - startX), static_cast<double>(py - startY)) >
                  radius))
                pts.emplace_back(px, py);
        }
    return pts;
}

// ----------------------------------------------------------------- legalizer
class Legalizer {
public:
    Legalizer(const sp_hip_design& d, const float* netW, float* x, float* y, float* offx,
              uint8_t* flags, uint8_t* orient, const sp_mlg_config& cfg, float site,
              float rowH, sp_metrics* sink, const char* const* names)
        : d_(d), netW_(netW), x_(x), y_(y), offx_(offx), flags_(flags), orient_(orient),
          cfg_(cfg), sink_(sink), names_(names) {
        snap_ = SiteSnap{d.core_lx, d.core_ly, d.core_hx, d.core_hy, site, rowH > 0.f ? rowH : 1.f};
        core_ = Rect4{d.core_lx, d.core_ly, d.core_hx, d.core_hy};
    }

    sp_mlg_result run() {
        sp_mlg_result r{};
        for (int i = 0; i < d_.num_movable; ++i)
            if ((flags_[i] & F_MACRO) && !(flags_[i] & F_FIXED)) node_.push_back(i);
        r.num_macros = static_cast<int>(node_.size());
        if (node_.empty()) return r;
        setUp(r);
        const Clock::time_point clock = Clock::now();
        anneal(r);
        r.sa_ms = msSince(clock);
        report("sa");
        repairLadder(r);
        pullBack();
        report("postlegal");
        report("final");
        writeBack(r);
        if (sink_) check(sp_metrics_record_duration(sink_, "mLG", msSince(clock)));
        return r;
    }

private:
    int count() const { return static_cast<int>(node_.size()); }
    double shiftOf(int s, float px, float py) const {
        return std::hypot(static_cast<double>(px - x0_[s]), static_cast<double>(py - y0_[s]));
    }
    double shiftOf(int s) const { return shiftOf(s, pl_.x(s), pl_.y(s)); }
    void snapTo(int s, float& px, float& py) const {
        snap_.apply(pl_.halfW(s), pl_.halfH(s), px, py);
    }
    double cost(double ov, double wl, double disp) const {
        return 8.0 * ov / refOverlap_ + 1.0 * wl / refWl_ + delta_ * disp / refDisp_;
    }
    static void check(int st) {
        if (st == SP_HIP_ENONFINITE) throw NonFinite(g_lastError);
        if (st != SP_HIP_OK) throw InvalidArg(g_lastError);
    }
    void setNumber(const std::string& k, double v) {
        if (sink_) check(sp_metrics_set_number(sink_, k.c_str(), v));
    }

    // Geometry, snapped start, index self-check, costs, start temperature.
    void setUp(sp_mlg_result& r) {
        const int n = count();
        std::vector<float> w(n), h(n);
        x0_.resize(n);
        y0_.resize(n);
        flip_.assign(n, 0);
        for (int s = 0; s < n; ++s) {
            const int m = node_[s];
            w[s] = d_.node_w[m];
            h[s] = d_.node_h[m];
            x0_[s] = x_[m];
            y0_[s] = y_[m];
            if (!std::isfinite(x0_[s]) || !std::isfinite(y0_[s]))
                throw InvalidArg("macro legalization: non-finite macro position");
        }
        // fixed nodes of positive area are blockages (terminal_NI are not)
        std::vector<Rect4> blocks;
        for (int i = d_.num_movable; i < d_.num_nodes; ++i) {
            if (flags_[i] & F_NI) continue;
            if (d_.node_w[i] <= 0.f || d_.node_h[i] <= 0.f) continue;
            blocks.push_back(Rect4{x_[i] - 0.5f * d_.node_w[i], y_[i] - 0.5f * d_.node_h[i],
                                   x_[i] + 0.5f * d_.node_w[i], y_[i] + 0.5f * d_.node_h[i]});
        }
        pl_.build(w, h, x0_, y0_, std::move(blocks), core_);
        for (int s = 0; s < n; ++s) {
            float px = pl_.x(s), py = pl_.y(s);
            snapTo(s, px, py);
            pl_.place(s, px, py);
        }
        lanes_ = std::make_unique<Lanes>(std::clamp<int>(
            static_cast<int>(std::thread::hardware_concurrency()), 1, 16));
        probes_.resize(lanes_->lanes());
        if (n <= 8000) selfCheck();

        nets_.build(d_, netW_, offx_, node_, x_, y_);
        r.hpwl_before = designHpwl(d_, x_, y_, offx_);
        overlap_ = pl_.total(probes_[0]);
        r.overlap_before = overlap_;
        double macroArea = 0.0;
        for (int m : node_)
            macroArea += static_cast<double>(d_.node_w[m]) * static_cast<double>(d_.node_h[m]);
        wl_ = nets_.total(pl_.xs(), pl_.ys(), flip_);
        refLen_ = std::sqrt(std::max(macroArea / static_cast<double>(n), 1.0));
        refOverlap_ = std::max(macroArea, 1.0);
        refWl_ = std::max(wl_, 1.0);
        refDisp_ = refLen_ * static_cast<double>(n);
        delta_ = static_cast<double>(cfg_.disp_weight);
        disp_ = 0.0;
        for (int s = 0; s < n; ++s) disp_ += shiftOf(s);
    }

    // The index against an independent full scan (OverlapModel::selfCheck,
    // macro_sa.cpp:401-418): a disagreement above 1e-3 is an error.
    void selfCheck() {
        std::vector<double> dev(count(), 0.0);
        lanes_->run(count(), [&](int s, int lane) {
            dev[s] = std::fabs(pl_.scanAt(s) - pl_.overlapAt(s, probes_[lane]));
        });
        double worst = 0.0;
        for (double v : dev) worst = std::max(worst, v);
        if (!std::isfinite(worst) || worst > 1e-3)
            throw InvalidArg("macro legalization: spatial index disagrees with a full scan");
    }

    // The macros that overlap something now.
    void collectHot(std::vector<int>& hot) {
        hot.clear();
        for (int s = 0; s < count(); ++s)
            if (pl_.overlapAt(s, probes_[0]) > 0.0) hot.push_back(s);
    }

    // The annealing chain (macro_sa.cpp:556-796): the start temperature from
    // 2000 trial shifts, then the moves, the best state restored at the end.
    void anneal(sp_mlg_result& r) {
        const int n = count();
        Probe& pr = probes_[0];
        std::vector<int> hot;
        collectHot(hot);
        r.hot_macros = static_cast<int>(hot.size());
        const int moves = cfg_.moves > 0 ? cfg_.moves : std::max(20000, r.hot_macros * 300);
        Draws rng(cfg_.random_seed ^ 0x9E3779B97F4A7C15ull);
        std::vector<std::pair<int, double>> fresh, fresh2;

        // the median worsening of trial shifts sets the start temperature
        const double acceptStart =
            std::clamp(static_cast<double>(cfg_.init_accept), 1e-6, 1.0 - 1e-6);
        double median = 0.0;
        {
            std::vector<double> worse;
            worse.reserve(2000);
            const double here = cost(overlap_, wl_, disp_);
            const float reach = static_cast<float>(refLen_);
            for (int k = 0; k < 2000; ++k) {
                const int s = rng.pick(0, n - 1);
                const float ox = pl_.x(s), oy = pl_.y(s);
                float px = ox + rng.real(-reach, reach);
                float py = oy + rng.real(-reach, reach);
                snapTo(s, px, py);
                const double dOv = pl_.overlapAt(s, px, py, pr) - pl_.overlapAt(s, ox, oy, pr);
                const double dD = shiftOf(s, px, py) - shiftOf(s, ox, oy);
                pl_.place(s, px, py);
                const double dW = nets_.change(s, pl_.xs(), pl_.ys(), flip_, fresh);
                pl_.place(s, ox, oy);
                const double dc = cost(overlap_ + dOv, wl_ + dW, disp_ + dD) - here;
                if (dc > 0.0) worse.push_back(dc);
            }
            if (!worse.empty()) {
                const size_t mid = worse.size() / 2;
                std::nth_element(worse.begin(), worse.begin() + static_cast<long>(mid),
                                 worse.end());
                median = worse[mid];
            }
        }
        const double t0 = median > 0.0 ? -median / std::log(acceptStart)
                                       : -1.0 / std::log(acceptStart);
        const double ratio = std::max(1e-12f, cfg_.temp_ratio);
        r.t0 = t0;
        r.typical_worsen = median;

        std::vector<float> bestX = pl_.xs(), bestY = pl_.ys();
        std::vector<uint8_t> bestFlip = flip_;
        double bestCost = cost(overlap_, wl_, disp_);
        int lastGain = 0, done = 0;
        for (int it = 0; it < moves; ++it) {
            done = it + 1;
            const double progress = static_cast<double>(it) / static_cast<double>(moves);
            const double temp = t0 * std::pow(ratio, progress);
            if (it > 0 && it % 500 == 0) collectHot(hot);
            if (overlap_ <= 0.0 && it - lastGain > 20000) break;
            const float span = static_cast<float>(refLen_ * (0.5 * (1.0 - progress) + 0.02));

            const bool aimed = !hot.empty() && rng.real(0.f, 1.f) < 0.7f;
            const int s = aimed ? hot[rng.pick(0, static_cast<int>(hot.size()) - 1)]
                                : rng.pick(0, n - 1);
            const int kind = aimed ? 0 : rng.pick(0, 9);   // 0-6 shift, 7-8 swap, 9 flip
            const float sx = pl_.x(s), sy = pl_.y(s);
            const uint8_t sFlip = flip_[s];
            int other = -1;
            float tx = 0.f, ty = 0.f;
            double dOv = 0.0, dD = 0.0;
            if (kind <= 6) {
                float px = sx, py = sy;
                Rect4 b{};
                if (aimed && pl_.heaviest(s, sx, sy, b, pr)) {
                    // off the heaviest neighbour by the least shift on an axis
                    const Rect4 a = pl_.box(s, sx, sy);
                    const float dir[4][2] = {{b.lx - a.hx, 0.f}, {b.hx - a.lx, 0.f},
                                             {0.f, b.ly - a.hy}, {0.f, b.hy - a.ly}};
                    int k = 0;
                    if (rng.real(0.f, 1.f) < 0.7f) {
                        float least = kFloatInf;
                        for (int q = 0; q < 4; ++q) {
                            const float len = std::fabs(dir[q][0]) + std::fabs(dir[q][1]);
                            if (len < least) {
                                least = len;
                                k = q;
                            }
                        }
                    } else {
                        k = rng.pick(0, 3);
                    }
                    px = sx + dir[k][0];
                    py = sy + dir[k][1];
                } else {
                    px = sx + rng.real(-span, span);
                    py = sy + rng.real(-span, span);
                }
                snapTo(s, px, py);
                dOv = pl_.overlapAt(s, px, py, pr) - pl_.overlapAt(s, sx, sy, pr);
                dD = shiftOf(s, px, py) - shiftOf(s, sx, sy);
                pl_.place(s, px, py);
            } else if (kind <= 8 && n >= 2) {
                do {
                    other = rng.pick(0, n - 1);
                } while (other == s);
                tx = pl_.x(other);
                ty = pl_.y(other);
                // each macro's overlap counts the pair once too many
                const double before = pl_.overlapAt(s, sx, sy, pr) +
                                      pl_.overlapAt(other, tx, ty, pr) -
                                      cover(pl_.box(s, sx, sy), pl_.box(other, tx, ty));
                float ax = tx, ay = ty, bx = sx, by = sy;
                snapTo(s, ax, ay);
                snapTo(other, bx, by);
                pl_.place(s, ax, ay);
                pl_.place(other, bx, by);
                const double after = pl_.overlapAt(s, ax, ay, pr) +
                                     pl_.overlapAt(other, bx, by, pr) -
                                     cover(pl_.box(s, ax, ay), pl_.box(other, bx, by));
                dOv = after - before;
                dD = shiftOf(s, ax, ay) - shiftOf(s, sx, sy) + shiftOf(other, bx, by) -
                     shiftOf(other, tx, ty);
            } else {
                flip_[s] = static_cast<uint8_t>(1 - sFlip);   // pins only
            }
            double dW = nets_.change(s, pl_.xs(), pl_.ys(), flip_, fresh);
            fresh2.clear();
            if (other >= 0) dW += nets_.change(other, pl_.xs(), pl_.ys(), flip_, fresh2);
            const double dCost = cost(overlap_ + dOv, wl_ + dW, disp_ + dD) -
                                 cost(overlap_, wl_, disp_);
            const bool take =
                dCost <= 0.0 || (temp > 0.0 && rng.real(0.f, 1.f) < std::exp(-dCost / temp));
            if (!take) {
                pl_.place(s, sx, sy);
                flip_[s] = sFlip;
                if (other >= 0) pl_.place(other, tx, ty);
                continue;
            }
            overlap_ += dOv;
            wl_ += dW;
            disp_ += dD;
            nets_.accept(fresh);
            if (other >= 0) nets_.accept(fresh2);
            ++r.accepted;
            const double c = cost(overlap_, wl_, disp_);
            if (c < bestCost) {
                bestCost = c;
                bestX = pl_.xs();
                bestY = pl_.ys();
                bestFlip = flip_;
                lastGain = it;
            }
        }
        r.moves = done;
        for (int s = 0; s < n; ++s) pl_.place(s, bestX[s], bestY[s]);
        flip_ = bestFlip;
        wl_ = nets_.total(pl_.xs(), pl_.ys(), flip_);
        overlap_ = pl_.total(pr);
        disp_ = 0.0;
        for (int s = 0; s < n; ++s) disp_ += shiftOf(s);
        r.overlap_after_sa = overlap_;
    }

    // Per-step summary keys (mlg_<step>_*): overlap, mean and largest
    // displacement from the start, the macro with the largest one and where
    // it started and is.
    void report(const std::string& step) {
        double sum = 0.0, most = -1.0;
        int far = 0;
        for (int s = 0; s < count(); ++s) {
            const double v = shiftOf(s);
            sum += v;
            if (v > most) {
                most = v;
                far = s;
            }
        }
        if (!sink_) return;
        const std::string p = "mlg_" + step + "_";
        setNumber(p + "overlap", pl_.total(probes_[0]));
        setNumber(p + "disp_mean", sum / count());
        setNumber(p + "disp_max", most);
        const int node = node_[far];
        const std::string name = names_ ? std::string(names_[node]) : std::to_string(node);
        check(sp_metrics_set_string(sink_, (p + "worst_macro").c_str(), name.c_str()));
        setNumber(p + "original_x", x0_[far]);
        setNumber(p + "original_y", y0_[far]);
        setNumber(p + "x", pl_.x(far));
        setNumber(p + "y", pl_.y(far));
    }

    // Repairs from the smallest moves to the largest (macro_sa.cpp:1164-1181).
    void repairLadder(sp_mlg_result& r) {
        constexpr double kNearCap = 3.0;   // grid search within 3 macro sides first
        separation(r);
        report("separate1");
        chains(r);
        report("chain1");
        gridSearch(kNearCap, r);
        report("local_grid");
        separation(r);
        report("separate2");
        chains(r);
        report("chain2");
        if (overlap_ > 0.0) {
            gridSearch(kDoubleInf, r);
            report("unlimited_grid");
            separation(r);
            chains(r);
        }
    }

    // One pass set of the minimum-separation shift (macro_sa.cpp:963-1006):
    // each overlapping macro, in index order, leaves its heaviest neighbour
    // by the shortest of the four axis shifts that strictly lowers its
    // overlap; with `tieShorter`, among equal overlaps the smaller
    // displacement wins.  Stops when a pass moves nothing or does not lower
    // the total.
    void separationPasses(bool tieShorter, int& repaired) {
        Probe& pr = probes_[0];
        for (int pass = 0; pass < 30 && overlap_ > 0.0; ++pass) {
            bool any = false;
            for (int s = 0; s < count(); ++s) {
                const float cx = pl_.x(s), cy = pl_.y(s);
                if (pl_.overlapAt(s, pr) <= 0.0) continue;
                const Rect4 a = pl_.box(s);
                Rect4 b{};
                if (!pl_.heaviest(s, cx, cy, b, pr)) continue;
                const float shifts[4][2] = {{b.lx - a.hx, 0.f}, {b.hx - a.lx, 0.f},
                                            {0.f, b.ly - a.hy}, {0.f, b.hy - a.ly}};
                const double now = pl_.overlapAt(s, cx, cy, pr);
                double bestOv = now, bestShift = kDoubleInf;
                float bx = cx, by = cy;
                for (const auto& sh : shifts) {
                    float px = cx + sh[0], py = cy + sh[1];
                    snapTo(s, px, py);
                    const double o = pl_.overlapAt(s, px, py, pr);
                    const double moved = shiftOf(s, px, py);
                    const bool tie = tieShorter && std::fabs(o - bestOv) <= 1e-9 && moved < bestShift;
                    if (o < now - 1e-9 && (o < bestOv - 1e-9 || tie)) {
                        bestOv = o;
                        bestShift = moved;
                        bx = px;
                        by = py;
                    }
                }
                if (bx != cx || by != cy) {
                    pl_.place(s, bx, by);
                    ++repaired;
                    any = true;
                }
            }
            const double after = pl_.total(pr);
            const bool stop = !any || after >= overlap_;
            overlap_ = after;
            if (stop) break;
        }
    }

    // Both tie rules from the same start; the small-move result is kept
    // unless it leaves more overlap or a larger total or largest
    // displacement (macro_sa.cpp:1008-1043).
    void separation(sp_mlg_result& r) {
        if (overlap_ <= 0.0) return;
        const std::vector<float> inX = pl_.xs(), inY = pl_.ys();
        const double inOverlap = overlap_;
        int repaired = r.repaired;
        separationPasses(false, repaired);
        const std::vector<float> plainX = pl_.xs(), plainY = pl_.ys();
        const double plainOverlap = overlap_;
        const int plainRepaired = repaired;
        const auto spread = [&] {
            double sum = 0.0, most = 0.0;
            for (int s = 0; s < count(); ++s) {
                const double v = shiftOf(s);
                sum += v;
                most = std::max(most, v);
            }
            return std::make_pair(sum, most);
        };
        const auto plainSpread = spread();
        for (int s = 0; s < count(); ++s) pl_.place(s, inX[s], inY[s]);
        overlap_ = inOverlap;
        repaired = r.repaired;
        separationPasses(true, repaired);
        const auto tieSpread = spread();
        if (overlap_ > plainOverlap || tieSpread.first > plainSpread.first ||
            tieSpread.second > plainSpread.second) {
            for (int s = 0; s < count(); ++s) pl_.place(s, plainX[s], plainY[s]);
            overlap_ = plainOverlap;
            repaired = plainRepaired;
        }
        r.repaired = repaired;
    }

    // Chain pushes (macro_sa.cpp:1045-1162): a macro pushed along an axis
    // pushes the macros ahead of it by what they need, on grid units, up to
    // 64 macros and a reach of the first push plus two typical sides; the
    // whole chain moves as one if it lowers the chain's own overlap.
    struct Chain {
        std::vector<float> push;       // per macro, the push it needs (this pass)
        std::vector<int> seenAt;       // per macro, the build that enqueued it
        std::vector<int> members, hits;
        std::vector<float> keepX, keepY;
        int build = 0;
    };

    bool formChain(Chain& c, int s0, int axis, int sign, float first) {
        constexpr int kLongest = 64;
        const float unit = axis == 0 ? snap_.site : snap_.row;
        const auto toGrid = [unit](float v) { return std::ceil(v / unit - 1e-4f) * unit; };
        const float reach = first + 2.f * static_cast<float>(refLen_);
        ++c.build;
        c.members.assign(1, s0);
        c.seenAt[s0] = c.build;
        c.push[s0] = toGrid(first);
        for (size_t q = 0; q < c.members.size(); ++q) {
            if (q > 4 * kLongest) return false;
            const int t = c.members[q];
            const float step = static_cast<float>(sign) * c.push[t];
            const Rect4 a = pl_.box(t, pl_.x(t) + (axis == 0 ? step : 0.f),
                                    pl_.y(t) + (axis == 1 ? step : 0.f));
            if (a.lx < core_.lx - 1e-3f || a.hx > core_.hx + 1e-3f || a.ly < core_.ly - 1e-3f ||
                a.hy > core_.hy + 1e-3f)
                return false;
            pl_.macrosHitting(a, t, c.hits);
            const float at = axis == 0 ? pl_.x(t) : pl_.y(t);
            for (int u : c.hits) {
                const float au = axis == 0 ? pl_.x(u) : pl_.y(u);
                if ((au - at) * static_cast<float>(sign) <= 0.f) continue;   // behind
                const Rect4 b = pl_.box(u);
                const float gap = sign > 0 ? (axis == 0 ? a.hx - b.lx : a.hy - b.ly)
                                           : (axis == 0 ? b.hx - a.lx : b.hy - a.ly);
                const float need = toGrid(gap);
                if (need > reach) return false;
                const bool known = c.seenAt[u] == c.build;
                if (known && need <= c.push[u]) continue;
                if (!known) {
                    if (static_cast<int>(c.members.size()) >= kLongest) return false;
                    c.seenAt[u] = c.build;
                }
                c.push[u] = need;
                c.members.push_back(u);   // again when its push grew
            }
        }
        std::sort(c.members.begin(), c.members.end());
        c.members.erase(std::unique(c.members.begin(), c.members.end()), c.members.end());
        return true;
    }

    double chainOverlap(const Chain& c) {
        Probe& pr = probes_[0];
        double sum = 0.0;
        for (int u : c.members) sum += pl_.overlapAt(u, pr);
        for (size_t i = 0; i < c.members.size(); ++i)
            for (size_t j = i + 1; j < c.members.size(); ++j)
                sum -= cover(pl_.box(c.members[i]), pl_.box(c.members[j]));
        return sum;
    }

    // Moves the chain; returns how much its overlap fell, and puts it back
    // unless `keep`.
    double shiftChain(Chain& c, int axis, int sign, bool keep) {
        const double before = chainOverlap(c);
        const size_t m = c.members.size();
        c.keepX.resize(m);
        c.keepY.resize(m);
        for (size_t k = 0; k < m; ++k) {
            const int u = c.members[k];
            c.keepX[k] = pl_.x(u);
            c.keepY[k] = pl_.y(u);
            const float step = static_cast<float>(sign) * c.push[u];
            pl_.place(u, c.keepX[k] + (axis == 0 ? step : 0.f), c.keepY[k] + (axis == 1 ? step : 0.f));
        }
        const double gain = before - chainOverlap(c);
        if (!keep)
            for (size_t k = 0; k < m; ++k) pl_.place(c.members[k], c.keepX[k], c.keepY[k]);
        return gain;
    }

    void chains(sp_mlg_result& r) {
        Probe& pr = probes_[0];
        for (int pass = 0; pass < 20 && overlap_ > 0.0; ++pass) {
            Chain c;
            c.push.assign(count(), 0.f);
            c.seenAt.assign(count(), -1);
            bool any = false;
            for (int s = 0; s < count(); ++s) {
                if (pl_.overlapAt(s, pr) <= 0.0) continue;
                Rect4 b{};
                if (!pl_.heaviest(s, pl_.x(s), pl_.y(s), b, pr)) continue;
                const Rect4 a = pl_.box(s);
                // left, right, down, up: the push that clears b
                const float amount[4] = {a.hx - b.lx, b.hx - a.lx, a.hy - b.ly, b.hy - a.ly};
                int pick = -1;
                double bestGain = 1e-6;
                for (int k = 0; k < 4; ++k) {
                    const int axis = k / 2, sign = (k % 2) ? 1 : -1;
                    if (!(amount[k] > 0.f) || !formChain(c, s, axis, sign, amount[k])) continue;
                    const double g = shiftChain(c, axis, sign, false);
                    if (g > bestGain) {
                        bestGain = g;
                        pick = k;
                    }
                }
                if (pick < 0) continue;
                const int axis = pick / 2, sign = (pick % 2) ? 1 : -1;
                formChain(c, s, axis, sign, amount[pick]);
                shiftChain(c, axis, sign, true);
                r.chained += static_cast<int>(c.members.size());
                any = true;
            }
            const double after = pl_.total(pr);
            const bool stop = !any || after >= overlap_;
            overlap_ = after;
            if (stop) break;
        }
    }

    // The obstacles a contact search of macro s may reach: blockages, then
    // the other macros in index order, as the reference lists them (only
    // those within reach survive its filter, and the point lists are sorted,
    // so nothing farther is needed).
    std::vector<Rect4> obstaclesAround(int s, double radius) const {
        std::vector<Rect4> out;
        const float hw = pl_.width(s) * .5f, hh = pl_.height(s) * .5f;
        const double x = x0_[s], y = y0_[s];
        const double pad = radius + std::max(hw, hh) + 1.0;
        const bool bounded = std::isfinite(pad);
        auto near = [&](const Rect4& b) {
            return !bounded || !(b.hx < x - pad || b.lx > x + pad || b.hy < y - pad || b.ly > y + pad);
        };
        for (const Rect4& b : pl_.blockages())
            if (near(b)) out.push_back(b);
        for (int t = 0; t < count(); ++t)
            if (t != s) {
                const Rect4 b = pl_.box(t);
                if (near(b)) out.push_back(b);
            }
        return out;
    }

    // Grid search (macro_sa.cpp:822-924): each overlapping macro, heaviest
    // first, tries a (2 span + 1)^2 grid of positions around where it is
    // (steps of a quarter side, at least a site / a row) and the contact
    // points around its start, and takes the zero-overlap position nearest
    // its start within zeroCap sides; failing that, the grid position of
    // least overlap within the pass's displacement limit.
    void gridSearch(double zeroCap, sp_mlg_result& r) {
        const int n = count();
        std::vector<int> order(n);
        for (int s = 0; s < n; ++s) order[s] = s;
        struct Cand {
            float x, y;
            double shift, ov;
        };
        std::vector<Cand> cand;
        for (int pass = 0; pass < 12 && overlap_ > 0.0; ++pass) {
            std::vector<double> ov(n);
            for (int s = 0; s < n; ++s) ov[s] = pl_.overlapAt(s, probes_[0]);
            std::stable_sort(order.begin(), order.end(),
                             [&ov](int a, int b) { return ov[a] > ov[b]; });
            for (int s : order) {
                const double now = pl_.overlapAt(s, probes_[0]);
                if (now <= 0.0) continue;
                const float cx = pl_.x(s), cy = pl_.y(s);
                const float stepX = std::max(pl_.width(s) * 0.25f, snap_.site);
                const float stepY = std::max(pl_.height(s) * 0.25f, snap_.row);
                const int span = 12 + 2 * pass, side1 = 2 * span + 1;
                const double side = std::max(pl_.width(s), pl_.height(s));
                const double zeroLimit = zeroCap * side;
                const double fbLimit = std::min(static_cast<double>(2 + pass) * side, zeroLimit);
                // grid positions, row by row (y outer), evaluated in parallel
                cand.resize(static_cast<size_t>(side1) * side1);
                lanes_->run(side1 * side1, [&](int i, int lane) {
                    const int ry = i / side1 - span, rx = i % side1 - span;
                    float px = cx + static_cast<float>(rx) * stepX;
                    float py = cy + static_cast<float>(ry) * stepY;
                    snapTo(s, px, py);
                    cand[i] = Cand{px, py, shiftOf(s, px, py), pl_.overlapAt(s, px, py, probes_[lane])};
                });
                bool zeroFound = false, fbFound = false;
                float zx = 0.f, zy = 0.f, fx = 0.f, fy = 0.f;
                double zeroShift = kDoubleInf, fbOv = now;
                for (const Cand& c : cand) {
                    if (c.ov <= 0.0 && c.shift <= zeroLimit) {
                        if (c.shift < zeroShift) {
                            zeroShift = c.shift;
                            zx = c.x;
                            zy = c.y;
                            zeroFound = true;
                        }
                    } else if (!zeroFound && c.shift <= fbLimit && c.ov < fbOv - 1e-9) {
                        fbOv = c.ov;
                        fx = c.x;
                        fy = c.y;
                        fbFound = true;
                    }
                }
                // contact points (and their site / row neighbours): only
                // zero-overlap positions nearer the start count
                const double radius = std::min(zeroLimit, static_cast<double>(2 + pass) * side);
                const auto pts = contactPoints(x0_[s], y0_[s], cx, cy, pl_.width(s),
                                               pl_.height(s), core_, obstaclesAround(s, radius),
                                               radius);
                cand.resize(pts.size() * 9);
                const double bound = zeroShift;
                lanes_->run(static_cast<int>(cand.size()), [&](int i, int lane) {
                    const auto& p = pts[i / 9];
                    const int dx = (i % 9) / 3 - 1, dy = i % 3 - 1;
                    float px = p.first + dx * snap_.site, py = p.second + dy * snap_.row;
                    snapTo(s, px, py);
                    const double sh = shiftOf(s, px, py);
                    const bool useful = sh <= zeroLimit && sh < bound;
                    cand[i] = Cand{px, py, sh,
                                   useful ? pl_.overlapAt(s, px, py, probes_[lane]) : 1.0};
                });
                for (const Cand& c : cand)
                    if (c.shift <= zeroLimit && c.shift < zeroShift && c.ov <= 0.0) {
                        zeroShift = c.shift;
                        zx = c.x;
                        zy = c.y;
                        zeroFound = true;
                    }
                if (zeroFound) {
                    pl_.place(s, zx, zy);
                    ++r.repaired;
                } else if (fbFound) {
                    pl_.place(s, fx, fy);
                    ++r.repaired;
                }
            }
            overlap_ = pl_.total(probes_[0]);
        }
    }

    // Once legal, the macro farthest from its start moves to the nearest
    // zero-overlap contact point (or neighbour) if that is nearer and does
    // not lengthen its nets (macro_sa.cpp:1182-1219).
    void pullBack() {
        if (overlap_ > 0.0) return;
        int far = 0;
        double farShift = 0.0;
        for (int s = 0; s < count(); ++s) {
            const double v = shiftOf(s);
            if (v > farShift) {
                farShift = v;
                far = s;
            }
        }
        const auto pts = contactPoints(x0_[far], y0_[far], pl_.x(far), pl_.y(far), pl_.width(far),
                                       pl_.height(far), core_, obstaclesAround(far, farShift),
                                       farShift);
        float bx = pl_.x(far), by = pl_.y(far);
        double best = farShift;
        for (const auto& p : pts)
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    float px = p.first + dx * snap_.site, py = p.second + dy * snap_.row;
                    snapTo(far, px, py);
                    const double v = shiftOf(far, px, py);
                    if (v < best && pl_.overlapAt(far, px, py, probes_[0]) <= 0.0) {
                        best = v;
                        bx = px;
                        by = py;
                    }
                }
        if (!(best < farShift)) return;
        nets_.total(pl_.xs(), pl_.ys(), flip_);   // the repairs moved macros since the chain
        const float ox = pl_.x(far), oy = pl_.y(far);
        pl_.place(far, bx, by);
        std::vector<std::pair<int, double>> fresh;
        const double dW = nets_.change(far, pl_.xs(), pl_.ys(), flip_, fresh);
        if (dW > 0.0) pl_.place(far, ox, oy);
        setNumber("mlg_postlegal_accepted", dW <= 0.0 ? 1 : 0);
    }

    // Positions, mirrored pins and orientations of flipped macros (N <-> FN,
    // S <-> FS), macros fixed; the final overlap and HPWL.
    void writeBack(sp_mlg_result& r) {
        static const uint8_t kMirror[8] = {4, 1, 6, 3, 0, 5, 2, 7};
        for (int s = 0; s < count(); ++s) {
            const int m = node_[s];
            x_[m] = pl_.x(s);
            y_[m] = pl_.y(s);
            if (flip_[s]) {
                for (int k = d_.node2pin_start[m]; k < d_.node2pin_start[m + 1]; ++k) {
                    const int p = d_.flat_node2pin[k];
                    offx_[p] = -offx_[p];
                }
                if (orient_ && orient_[m] < 8) orient_[m] = kMirror[orient_[m]];
            }
            flags_[m] |= F_FIXED;
        }
        r.overlap_after = pl_.total(probes_[0]);
        r.hpwl_after = designHpwl(d_, x_, y_, offx_);
    }

    const sp_hip_design& d_;
    const float* netW_;
    float *x_, *y_, *offx_;
    uint8_t *flags_, *orient_;
    sp_mlg_config cfg_;
    sp_metrics* sink_;
    const char* const* names_;
    SiteSnap snap_{};
    Rect4 core_{};
    std::vector<int> node_;            // design node of each macro slot
    std::vector<float> x0_, y0_;       // positions before legalization
    std::vector<uint8_t> flip_;
    Placement pl_;
    NetModel nets_;
    std::unique_ptr<Lanes> lanes_;
    std::vector<Probe> probes_;        // one per lane
    double overlap_ = 0.0, wl_ = 0.0, disp_ = 0.0;
    double refLen_ = 1.0, refOverlap_ = 1.0, refWl_ = 1.0, refDisp_ = 1.0, delta_ = 1.0;
};

}  // namespace
}  // namespace sphip

extern "C" int sp_macro_legalize(const sp_hip_design* d, const float* net_weight,
                                 float site_width, float row_height, const sp_mlg_config* cfg,
                                 float* x, float* y, uint8_t* node_flags, float* pin_offset_x,
                                 uint8_t* orient, sp_metrics* sink, const char* const* node_names,
                                 sp_mlg_result* result) {
    return sphip::guarded([&] {
        using sphip::InvalidArg;
        if (!d || !cfg || !x || !y || !node_flags || !pin_offset_x || !result)
            throw InvalidArg("null argument");
        if (!(site_width > 0.f)) throw InvalidArg("site width must be positive");
        if (d->num_pins > 0 && (!d->pin2node || !d->pin2net || !d->pin_offset_y ||
                                !d->flat_net2pin || !d->flat_node2pin))
            throw InvalidArg("null pin array");
        if (!d->net2pin_start || !d->node2pin_start) throw InvalidArg("null CSR array");
        sphip::Legalizer lg(*d, net_weight, x, y, pin_offset_x, node_flags, orient, *cfg,
                            site_width, row_height, sink, node_names);
        *result = lg.run();
    });
}
