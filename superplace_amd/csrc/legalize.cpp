// Standard-cell legalization by Abacus (P. Spindler, U. Schlichtmann,
// F. M. Johannes, "Abacus: fast legalization of standard cell circuits with
// minimal movement", ISPD 2008), as the reference runs it
// (legalizeCells / buildAbacusSegments, src/lg/abacus.cpp:108-286).  Host
// code: the algorithm is a sequential greedy pass over the cells.
//
// The result must equal the reference's bit for bit on the same input, so
// the steps and their arithmetic follow it:
//  * segments: every site row inside the core, minus the positive-area
//    intersections with fixed obstacles, shrunk to whole sites, kept if at
//    least as wide as the narrowest movable cell (double arithmetic);
//  * cells (movable, not macro / fixed / filler / NI, positive size) in
//    order of their lower-left x (index breaks ties), each tried on rows in
//    order of |row y - cell y| (below first on ties), stopping once that
//    distance squared reaches the best cost; per row the segment with room
//    whose clamped x is nearest; the trial appends the cell to the
//    segment's last cluster chain (Abacus clusters: e, q, w, x) without
//    changing it; cost = dx^2 + dy^2 of the cell's trial position;
//  * widths rounded up to whole sites (within 1e-5 of a whole count: that
//    count); cluster positions snapped to the nearest site, clamped into the
//    segment and kept right of the previous cluster;
//  * self-checks: overlaps, off-site or off-row cells, cells past the
//    segment's end; mean and max displacement of the placed cells.
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>
#include <vector>

#include "abi.h"

namespace sphip {
namespace {

struct Seg {
    int row;
    float lx, hx;
};

// Abacus cluster: cells first..last of a segment, weight e, q = sum of
// weight * (desired x - offset in the cluster), width w, position x.
struct Clu {
    int first = 0, last = 0;
    double e = 0.0, q = 0.0, w = 0.0, x = 0.0;
};

struct SegState {
    Seg seg{};
    double used = 0.0;
    std::vector<int> cells;
    std::vector<Clu> chain;
};

class Abacus {
public:
    Abacus(const sp_hip_design& d, const float* rowStep, float* x, float* y)
        : d_(d), step_(rowStep), x_(x), y_(y) {}

    float llx(int i) const { return x_[i] - 0.5f * d_.node_w[i]; }
    float lly(int i) const { return y_[i] - 0.5f * d_.node_h[i]; }
    float urx(int i) const { return x_[i] + 0.5f * d_.node_w[i]; }
    float ury(int i) const { return y_[i] + 0.5f * d_.node_h[i]; }
    float rowHy(int r) const { return d_.row_ly[r] + d_.row_height[r]; }
    bool flag(int i, uint8_t f) const { return (d_.node_flags[i] & f) != 0; }

    bool isCell(int i) const {
        return i < d_.num_movable && !flag(i, F_MACRO) && !flag(i, F_FIXED) &&
               !flag(i, F_FILLER) && !flag(i, F_NI) && d_.node_w[i] > 0.f && d_.node_h[i] > 0.f;
    }

    static double sitesWidth(float width, float step) {
        const double n = static_cast<double>(width) / step;
        const double k = std::round(n);
        return (std::abs(n - k) < 1e-5 ? k : std::ceil(n)) * step;
    }

    std::vector<Seg> segments() const {
        float minW = std::numeric_limits<float>::infinity();
        for (int i = 0; i < d_.num_movable; ++i)
            if (isCell(i)) minW = std::min(minW, d_.node_w[i]);
        if (!std::isfinite(minW)) minW = 0.f;
        std::vector<int> blockers;
        for (int i = 0; i < d_.num_nodes; ++i)
            if (flag(i, F_FIXED) && !flag(i, F_NI) && !flag(i, F_FILLER) && d_.node_w[i] > 0.f &&
                d_.node_h[i] > 0.f)
                blockers.push_back(i);
        std::vector<Seg> out;
        for (int r = 0; r < d_.num_rows; ++r) {
            const float ly = d_.row_ly[r], h = d_.row_height[r], lx = d_.row_lx[r];
            const float st = step_[r], hx = d_.row_hx[r];
            if (!std::isfinite(lx) || !std::isfinite(ly) || !std::isfinite(st) ||
                !std::isfinite(h) || st <= 0.f || h <= 0.f || !(hx > lx))
                throw InvalidArg("invalid site row for Abacus");
            if (ly < d_.core_ly || rowHy(r) > d_.core_hy) continue;
            const double low = std::max(lx, d_.core_lx), high = std::min(hx, d_.core_hx);
            if (low >= high) continue;
            std::vector<std::pair<double, double>> cut;
            for (int i : blockers) {
                if (lly(i) >= rowHy(r) || ury(i) <= ly) continue;
                const double a = std::max(low, static_cast<double>(llx(i)));
                const double b = std::min(high, static_cast<double>(urx(i)));
                if (a < b) cut.emplace_back(a, b);
            }
            std::sort(cut.begin(), cut.end());
            auto piece = [&](double a, double b) {
                if (a >= b) return;
                const double sl = lx + std::ceil((a - lx) / st) * st;
                const double sh = lx + std::floor((b - lx) / st) * st;
                if (sh > sl && sh - sl >= minW)
                    out.push_back({r, static_cast<float>(sl), static_cast<float>(sh)});
            };
            double at = low;
            for (const auto& c : cut) {
                piece(at, c.first);
                at = std::max(at, c.second);
            }
            piece(at, high);
        }
        return out;
    }

    static void place(Clu& c, const Seg& s) {
        c.x = std::clamp(c.q / c.e, static_cast<double>(s.lx), static_cast<double>(s.hx) - c.w);
    }
    static void append(Clu& c, int slot, double want, double width, double wt) {
        c.last = slot;
        c.e += wt;
        c.q += wt * (want - c.w);
        c.w += width;
    }
    static Clu merge(const Clu& a, const Clu& b) {
        Clu m = a;
        m.last = b.last;
        m.q += b.q - b.e * m.w;
        m.e += b.e;
        m.w += b.w;
        return m;
    }

    // The chain's new last cluster if the cell joined the segment; keep is
    // the number of leading clusters it leaves untouched.
    static Clu trial(const SegState& s, double want, double width, double wt, int& keep) {
        keep = static_cast<int>(s.chain.size());
        const int slot = static_cast<int>(s.cells.size());
        Clu t;
        if (keep && s.chain.back().x + s.chain.back().w > want) {
            t = s.chain.back();
            --keep;
        } else {
            t.first = slot;
        }
        append(t, slot, want, width, wt);
        place(t, s.seg);
        while (keep && s.chain[keep - 1].x + s.chain[keep - 1].w > t.x) {
            t = merge(s.chain[--keep], t);
            place(t, s.seg);
        }
        return t;
    }

    // The same step on a full copy of the chain (the reference's slow path).
    static Clu trialFull(const SegState& s, double want, double width, double wt,
                         std::vector<Clu>* into) {
        std::vector<Clu> ch = s.chain;
        const int slot = static_cast<int>(s.cells.size());
        if (ch.empty() || ch.back().x + ch.back().w <= want) {
            Clu f;
            f.first = slot;
            ch.push_back(f);
        }
        append(ch.back(), slot, want, width, wt);
        place(ch.back(), s.seg);
        while (ch.size() > 1) {
            const size_t l = ch.size() - 1;
            if (ch[l - 1].x + ch[l - 1].w <= ch[l].x) break;
            ch[l - 1] = merge(ch[l - 1], ch[l]);
            ch.pop_back();
            place(ch.back(), s.seg);
        }
        if (into) *into = ch;
        return ch.back();
    }

    sp_abacus_result run(bool fullTrials, bool areaWeight) {
        sp_abacus_result res{};
        res.precondition_ok = 1;
        res.rows = d_.num_rows;
        for (int i = 0; i < d_.num_movable; ++i)
            if (flag(i, F_MACRO) && !flag(i, F_FIXED)) {
                res.precondition_ok = 0;
                return res;
            }
        std::vector<int> order;
        for (int i = 0; i < d_.num_nodes; ++i) {
            if (!std::isfinite(x_[i]) || !std::isfinite(y_[i]) || !std::isfinite(d_.node_w[i]) ||
                !std::isfinite(d_.node_h[i]))
                throw InvalidArg("non-finite node geometry for Abacus");
            if (isCell(i)) order.push_back(i);
        }
        res.cells = static_cast<int>(order.size());
        if (order.empty()) return res;

        std::vector<float> x0(d_.num_movable), y0(d_.num_movable);
        for (int i : order) {
            x0[i] = llx(i);
            y0[i] = lly(i);
        }
        std::sort(order.begin(), order.end(),
                  [&](int a, int b) { return x0[a] < x0[b] || (x0[a] == x0[b] && a < b); });
        const std::vector<Seg> segs = segments();
        res.segments = static_cast<int>(segs.size());
        std::vector<SegState> st(segs.size());
        std::vector<std::vector<int>> ofRow(d_.num_rows);
        for (size_t s = 0; s < segs.size(); ++s) {
            st[s].seg = segs[s];
            ofRow[segs[s].row].push_back(static_cast<int>(s));
        }
        std::vector<int> rows(d_.num_rows);
        for (int r = 0; r < d_.num_rows; ++r) rows[r] = r;
        std::sort(rows.begin(), rows.end(), [&](int a, int b) {
            return d_.row_ly[a] < d_.row_ly[b] || (d_.row_ly[a] == d_.row_ly[b] && a < b);
        });
        const double inf = std::numeric_limits<double>::infinity();
        const int nr = d_.num_rows;

        for (int i : order) {
            const double wx = x0[i], wy = y0[i];
            const double wt = areaWeight ? static_cast<double>(d_.node_w[i]) * d_.node_h[i] : 1.0;
            int best = -1;
            double bestCost = inf;
            int up = static_cast<int>(std::lower_bound(rows.begin(), rows.end(), wy,
                                                       [&](int r, double v) { return d_.row_ly[r] < v; }) -
                                      rows.begin());
            int down = up - 1;
            while (down >= 0 || up < nr) {
                const double dDown = down >= 0 ? std::abs(d_.row_ly[rows[down]] - wy) : inf;
                const double dUp = up < nr ? std::abs(d_.row_ly[rows[up]] - wy) : inf;
                const double dist = std::min(dDown, dUp);
                if (dist * dist >= bestCost) break;
                const int r = dDown <= dUp ? rows[down--] : rows[up++];
                if (d_.node_h[i] > d_.row_height[r] + 1e-5f) continue;
                const double width = sitesWidth(d_.node_w[i], step_[r]);
                int cand = -1;
                double nearest = inf;
                for (int s : ofRow[r]) {
                    const SegState& ss = st[s];
                    if (ss.used + width > static_cast<double>(ss.seg.hx) - ss.seg.lx + 1e-7) continue;
                    const double dx = std::abs(std::clamp(wx, static_cast<double>(ss.seg.lx),
                                                          static_cast<double>(ss.seg.hx) - width) - wx);
                    if (dx < nearest) {
                        nearest = dx;
                        cand = s;
                    }
                }
                if (cand < 0) continue;
                int keep;
                const Clu t = fullTrials ? trialFull(st[cand], wx, width, wt, nullptr)
                                         : trial(st[cand], wx, width, wt, keep);
                const double dx = t.x + t.w - width - wx;
                const double cost = dx * dx + dist * dist;
                if (cost < bestCost) {
                    bestCost = cost;
                    best = cand;
                }
            }
            if (best < 0) {
                ++res.failed;
                continue;
            }
            SegState& ss = st[best];
            const float rstep = step_[ss.seg.row];
            const double width = sitesWidth(d_.node_w[i], rstep);
            if (std::abs(width - d_.node_w[i]) > 1e-5) ++res.rounded_widths;
            if (fullTrials) {
                trialFull(ss, wx, width, wt, &ss.chain);
            } else {
                int keep;
                const Clu t = trial(ss, wx, width, wt, keep);
                ss.chain.resize(keep);
                ss.chain.push_back(t);
            }
            ss.cells.push_back(i);
            ss.used += width;
        }

        double dispSum = 0.0;
        for (SegState& ss : st) {
            const int r = ss.seg.row;
            const float rlx = d_.row_lx[r], rly = d_.row_ly[r], rst = step_[r];
            double prevRight = ss.seg.lx;
            for (const Clu& c : ss.chain) {
                double px = rlx + std::round((c.x - rlx) / rst) * rst;
                px = std::clamp(px, static_cast<double>(ss.seg.lx), static_cast<double>(ss.seg.hx) - c.w);
                px = std::max(px, prevRight);
                for (int slot = c.first; slot <= c.last; ++slot) {
                    const int i = ss.cells[slot];
                    x_[i] = static_cast<float>(px) + 0.5f * d_.node_w[i];
                    y_[i] = rly + 0.5f * d_.node_h[i];
                    const double disp = std::hypot(static_cast<double>(llx(i)) - x0[i],
                                                   static_cast<double>(lly(i)) - y0[i]);
                    dispSum += disp;
                    res.disp_max = std::max(res.disp_max, disp);
                    px += sitesWidth(d_.node_w[i], rst);
                }
                prevRight = px;
            }
            float right = ss.seg.lx;
            for (int i : ss.cells) {
                if (llx(i) < right - 1e-3f) ++res.selfcheck_overlaps;
                const double sites = (static_cast<double>(llx(i)) - rlx) / rst;
                if (std::abs(sites - std::round(sites)) * rst > 1e-3 || std::abs(lly(i) - rly) > 1e-3 ||
                    urx(i) > ss.seg.hx + 1e-3f)
                    ++res.selfcheck_offsite;
                right = urx(i);
            }
        }
        const int placed = res.cells - res.failed;
        res.disp_mean = placed ? dispSum / placed : 0.0;
        return res;
    }

private:
    const sp_hip_design& d_;
    const float* step_;
    float* x_;
    float* y_;
};

}  // namespace
}  // namespace sphip

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_abacus_legalize(const sp_hip_design* d, const float* row_step, float* x, float* y,
                       int flags, sp_abacus_result* result) {
    return guarded([&] {
        if (!d || !row_step || !x || !y || !result) throw InvalidArg("null argument");
        if (d->num_rows > 0 && (!d->row_ly || !d->row_height || !d->row_lx || !d->row_hx))
            throw InvalidArg("null row array");
        if (d->num_nodes > 0 && (!d->node_w || !d->node_h || !d->node_flags))
            throw InvalidArg("null node array");
        sphip::Abacus a(*d, row_step, x, y);
        *result = a.run((flags & SP_ABACUS_FULL_TRIALS) != 0, (flags & SP_ABACUS_AREA_WEIGHT) != 0);
    });
}

int sp_abacus_segments(const sp_hip_design* d, const float* row_step, int32_t* seg_row,
                       float* seg_lx, float* seg_hx, int32_t cap, int32_t* count) {
    return guarded([&] {
        if (!d || !row_step || !count) throw InvalidArg("null argument");
        std::vector<float> x(d->node_x, d->node_x + d->num_nodes), y(d->node_y, d->node_y + d->num_nodes);
        sphip::Abacus a(*d, row_step, x.data(), y.data());
        const auto segs = a.segments();
        *count = static_cast<int32_t>(segs.size());
        if (!seg_row) return;
        if (cap < *count || !seg_lx || !seg_hx) throw InvalidArg("segment buffer too small");
        for (size_t s = 0; s < segs.size(); ++s) {
            seg_row[s] = segs[s].row;
            seg_lx[s] = segs[s].lx;
            seg_hx[s] = segs[s].hx;
        }
    });
}

}  // extern "C"
