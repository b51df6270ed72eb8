// Per-bin end of the density scatter, shared by the finalize kernel and by
// the first Poisson pass, which takes it over in the device loop (reading
// the integer grids instead of a finished rho): a bin's node and filler
// densities are its atomic-grid value plus every staged tile window that
// covers it (all integers, so the sum is exact); rho = (node + filler +
// terminal + base) / binArea (bin_grid.cpp:361-368) and the tau term
// max(0, (node + terminal + base)/area - target) area (bin_grid.cpp:373-393,
// fillers excluded).  The atomic grids are cleared for the next evaluation.
#pragma once

#include "gp_kernels.h"

namespace sphip {

template <typename T>
struct DensityIn {
    GridGeom<T> g;
    TileGrid tg;
    unsigned long long* nodeQ;
    unsigned long long* fillQ;
    const double* terminal;
    const double* base;
    double invQ, binArea, rhoShift;
    double* partials;    // one tau partial per block
    long long* dbgNode;  // optional: the integer sums per bin
    long long* dbgFill;
};

// Bin b = bx * dim + by: returns rho (shifted by rhoShift) and adds the tau
// term to *over.
template <typename T>
__device__ inline T finalizeBin(const DensityIn<T>& in, int b, double* over) {
    constexpr int WIN2 = kTileWin * kTileWin;
    const GridGeom<T>& g = in.g;
    const TileGrid& tg = in.tg;
    const int bx = b / g.dim, by = b - bx * g.dim;
    unsigned long long sn = in.nodeQ[b], sf = in.fillQ[b];
    in.nodeQ[b] = 0;
    in.fillQ[b] = 0;
    const int tx = bx / kTileBins, ty = by / kTileBins;
    const int ux = bx - tx * kTileBins + kTileHalo, uy = by - ty * kTileBins + kTileHalo;
    for (int dx = -1; dx <= 1; ++dx) {
        const int sx = ux - dx * kTileBins;
        if (sx < 0 || sx >= kTileWin || tx + dx < 0 || tx + dx >= tg.tilesX) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int sy = uy - dy * kTileBins;
            if (sy < 0 || sy >= kTileWin || ty + dy < 0 || ty + dy >= tg.tilesY) continue;
            const long at = static_cast<long>((tx + dx) * tg.tilesY + ty + dy) * 2 * WIN2 +
                            sx * kTileWin + sy;
            sn += tg.stage[at];
            sf += tg.stage[at + WIN2];
        }
    }
    const long long qn = static_cast<long long>(sn);
    const long long qf = static_cast<long long>(sf);
    const double node = static_cast<double>(qn) * in.invQ;
    const double fill = static_cast<double>(qf) * in.invQ;
    const double tb = in.terminal[b], bb = in.base[b];
    const double o = (node + tb + bb) / in.binArea - static_cast<double>(g.target);
    *over += o > 0.0 ? o * in.binArea : 0.0;
    if (in.dbgNode) in.dbgNode[b] = qn;
    if (in.dbgFill) in.dbgFill[b] = qf;
    // a constant shift only changes the DC term, which the solve discards;
    // it keeps the transform inputs small (smaller FP32 rounding)
    return static_cast<T>((node + fill + tb + bb) / in.binArea - in.rhoShift);
}

// The tile counters of the spatial sort, cleared with the grids (thread t of
// the pass clears key t).
__device__ inline void clearTileKeys(const TileGrid& tg, int t) {
    if (t < tg.nKeys) {
        tg.count[t] = 0;
        tg.fill[t] = 0;
    }
}

}  // namespace sphip
