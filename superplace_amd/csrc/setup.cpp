// Set-up either side of the placement loop, host code run once per design:
// the filler plan and the random filler scatter.
//
// Results must equal the reference's bit for bit on the same design, so the
// arithmetic follows initializeFillers / scatterFillers
// (src/gp/filler.cpp:33-139) and PlaceDB::computeRegions
// (src/db/place_db.cpp:97-120) operation by operation:
//  * whitespace = row area (double sum of float row extents) minus the
//    float overlap of every fixed, non-NI node with the rows, taken in row
//    order with the same early exit (rows ascend in y);
//  * node area scaled = cell area + target x macro area (double);
//  * filler area = mean area of the middle 80% of movable nodes, height the
//    most frequent row height, count = round(filler area / single area);
//  * filler centres: std::mt19937 seeded with the seed, one
//    std::uniform_real_distribution<float> draw for x then one for y per
//    filler, over the core shrunk by half the filler size.
#include <algorithm>
#include <cmath>
#include <map>
#include <random>
#include <vector>

#include "abi.h"

namespace sphip {
namespace {

struct Box {
    float lx, ly, hx, hy;
};

inline float overlap(const Box& a, const Box& b) {
    const float dx = std::min(a.hx, b.hx) - std::max(a.lx, b.lx);
    const float dy = std::min(a.hy, b.hy) - std::max(a.ly, b.ly);
    return (dx > 0.f && dy > 0.f) ? dx * dy : 0.f;
}

void checkDesign(const sp_hip_design* d) {
    if (!d) throw InvalidArg("null design");
    if (d->num_movable < 0 || d->num_nodes < d->num_movable || d->num_rows < 0)
        throw InvalidArg("bad node or row counts");
    if (d->num_nodes > 0 && (!d->node_x || !d->node_y || !d->node_w || !d->node_h ||
                             !d->node_flags))
        throw InvalidArg("null node array");
    if (d->num_rows > 0 && (!d->row_ly || !d->row_height || !d->row_lx || !d->row_hx))
        throw InvalidArg("null row array");
}

// Most frequent row height, ties to the smaller one; 0 without rows.
float modeRowHeight(const sp_hip_design* d) {
    std::map<float, int> count;
    for (int r = 0; r < d->num_rows; ++r) ++count[d->row_height[r]];
    float best = 0.f;
    int bestCount = -1;
    for (const auto& [h, c] : count)
        if (c > bestCount) {
            bestCount = c;
            best = h;
        }
    return best;
}

sp_filler_info fillerPlan(const sp_hip_design* d, float target) {
    sp_filler_info info{};
    const int rows = d->num_rows;
    double rowArea = 0.0;
    for (int r = 0; r < rows; ++r)
        rowArea += static_cast<double>(d->row_hx[r] - d->row_lx[r]) *
                   static_cast<double>(d->row_height[r]);

    double blocked = 0.0;
    for (int i = d->num_movable; i < d->num_nodes; ++i) {
        if (d->node_flags[i] & F_NI) continue;
        const float hw = 0.5f * d->node_w[i], hh = 0.5f * d->node_h[i];
        const Box b{d->node_x[i] - hw, d->node_y[i] - hh, d->node_x[i] + hw, d->node_y[i] + hh};
        for (int r = 0; r < rows; ++r) {
            const float ly = d->row_ly[r], hy = ly + d->row_height[r];
            if (ly >= b.hy) break;
            if (hy <= b.ly) continue;
            blocked += static_cast<double>(overlap(b, Box{d->row_lx[r], ly, d->row_hx[r], hy}));
        }
    }
    info.whitespace_area = std::max(0.0, rowArea - blocked);

    const int nMov = d->num_movable;
    std::vector<double> areas(static_cast<size_t>(nMov));
    double cells = 0.0, macros = 0.0;
    for (int i = 0; i < nMov; ++i) {
        areas[i] = static_cast<double>(d->node_w[i]) * static_cast<double>(d->node_h[i]);
        (d->node_flags[i] & F_MACRO ? macros : cells) += areas[i];
    }
    info.node_area_scaled = cells + macros * static_cast<double>(target);
    info.total_filler_area = info.whitespace_area * static_cast<double>(target) -
                             info.node_area_scaled;
    if (nMov == 0) return info;

    std::sort(areas.begin(), areas.end());
    const int lo = static_cast<int>(0.10 * nMov), hi = static_cast<int>(0.90 * nMov);
    double mean = 0.0;
    if (hi > lo) {
        for (int i = lo; i < hi; ++i) mean += areas[i];
        mean /= (hi - lo);
    } else {
        mean = areas[nMov / 2];
    }
    const float rh = modeRowHeight(d);
    info.height = rh > 0.f ? rh : 1.f;
    info.width = static_cast<float>(mean) / info.height;
    if (!(info.width > 0.f)) info.width = 1.f;
    const double single = static_cast<double>(info.width) * info.height;
    info.count = std::max(0, static_cast<int>(info.total_filler_area / std::max(single, 1e-9) + 0.5));
    return info;
}

void fillerScatter(const sp_hip_design* d, int count, float w, float h, uint64_t seed, float* x,
                   float* y) {
    std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
    const float hw = 0.5f * w, hh = 0.5f * h;
    const float loX = d->core_lx + hw, hiX = d->core_hx - hw;
    const float loY = d->core_ly + hh, hiY = d->core_hy - hh;
    const float cx = 0.5f * (d->core_lx + d->core_hx), cy = 0.5f * (d->core_ly + d->core_hy);
    for (int i = 0; i < count; ++i) {
        x[i] = hiX > loX ? std::uniform_real_distribution<float>(loX, hiX)(gen) : cx;
        y[i] = hiY > loY ? std::uniform_real_distribution<float>(loY, hiY)(gen) : cy;
    }
}

}  // namespace
}  // namespace sphip

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_filler_plan(const sp_hip_design* d, float target_density, sp_filler_info* info) {
    return guarded([&] {
        sphip::checkDesign(d);
        if (!info) throw InvalidArg("null argument");
        if (!(target_density > 0.f) || !std::isfinite(target_density))
            throw InvalidArg("target density must be positive");
        *info = sphip::fillerPlan(d, target_density);
    });
}

int sp_rng_uniform(uint64_t seed, int64_t count, float lo, float hi, float* out) {
    return guarded([&] {
        if (count < 0) throw InvalidArg("negative count");
        if (count > 0 && !out) throw InvalidArg("null output");
        if (!(lo < hi) || !std::isfinite(lo) || !std::isfinite(hi))
            throw InvalidArg("uniform range must be finite with lo < hi");
        std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
        std::uniform_real_distribution<float> dist(lo, hi);
        for (int64_t i = 0; i < count; ++i) out[i] = dist(gen);
    });
}

int sp_filler_scatter(const sp_hip_design* d, int count, float width, float height,
                      uint64_t seed, float* x, float* y) {
    return guarded([&] {
        sphip::checkDesign(d);
        if (count < 0) throw InvalidArg("negative filler count");
        if (count > 0 && (!x || !y)) throw InvalidArg("null output");
        sphip::fillerScatter(d, count, width, height, seed, x, y);
    });
}

}  // extern "C"
