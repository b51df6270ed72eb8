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

// The loads of bin b = bx * dim + by: its atomic-grid values, the staged
// windows that cover it (up to two tiles per axis) and its static terms.
// Split from finalizeFinish so that a thread can have the loads of several
// bins in flight before the clearing stores.
struct BinLoads {
    unsigned long long sn, sf;
    unsigned long long wn[4], wf[4];
    double tb, bb;
};

template <typename T>
__device__ inline void finalizeLoad(const DensityIn<T>& in, int b, BinLoads& L) {
    constexpr int WIN2 = kTileWin * kTileWin;
    const GridGeom<T>& g = in.g;
    const TileGrid& tg = in.tg;
    const int bx = b / g.dim, by = b - bx * g.dim;
    L.sn = in.nodeQ[b];
    L.sf = in.fillQ[b];
    L.tb = in.terminal[b];
    L.bb = in.base[b];
    // A window spans its tile and kTileHalo < kTileBins / 2 bins either side,
    // so per axis a bin lies in its own tile's window and at most one
    // neighbour's: the left one's when it sits in the first kTileHalo bins of
    // its tile, the right one's in the last kTileHalo.  Slots: own tile,
    // y-neighbour, x-neighbour, diagonal (fixed slots keep them in registers).
    static_assert(2 * kTileHalo < kTileBins, "a bin lies in at most two windows per axis");
    const int tx = bx / kTileBins, ty = by / kTileBins;
    const int ux = bx - tx * kTileBins, uy = by - ty * kTileBins;
    const int ndx = ux < kTileHalo ? -1 : (ux >= kTileBins - kTileHalo ? 1 : 0);
    const int ndy = uy < kTileHalo ? -1 : (uy >= kTileBins - kTileHalo ? 1 : 0);
    const bool okx = ndx != 0 && tx + ndx >= 0 && tx + ndx < tg.tilesX;
    const bool oky = ndy != 0 && ty + ndy >= 0 && ty + ndy < tg.tilesY;
    auto at = [&](int dx, int dy) {
        return static_cast<long>((tx + dx) * tg.tilesY + ty + dy) * 2 * WIN2 +
               (ux + kTileHalo - dx * kTileBins) * kTileWin + (uy + kTileHalo - dy * kTileBins);
    };
    const long a0 = at(0, 0);
    L.wn[0] = tg.stage[a0];
    L.wf[0] = tg.stage[a0 + WIN2];
    L.wn[1] = L.wf[1] = L.wn[2] = L.wf[2] = L.wn[3] = L.wf[3] = 0;
    if (oky) {
        const long a = at(0, ndy);
        L.wn[1] = tg.stage[a];
        L.wf[1] = tg.stage[a + WIN2];
    }
    if (okx) {
        const long a = at(ndx, 0);
        L.wn[2] = tg.stage[a];
        L.wf[2] = tg.stage[a + WIN2];
    }
    if (okx && oky) {
        const long a = at(ndx, ndy);
        L.wn[3] = tg.stage[a];
        L.wf[3] = tg.stage[a + WIN2];
    }
}

// A bin's node and filler densities from its loads (integer sums, so exact
// in any order); clears its atomic-grid entries; returns rho (shifted by
// rhoShift) and adds the tau term to *over.
template <typename T>
__device__ inline T finalizeFinish(const DensityIn<T>& in, int b, const BinLoads& L,
                                   double* over) {
    // only entries a spill wrote are cleared (most stay zero between sorts)
    if (L.sn) in.nodeQ[b] = 0;
    if (L.sf) in.fillQ[b] = 0;
    unsigned long long sn = L.sn, sf = L.sf;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        sn += L.wn[w];
        sf += L.wf[w];
    }
    const long long qn = static_cast<long long>(sn);
    const long long qf = static_cast<long long>(sf);
    const double node = static_cast<double>(qn) * in.invQ;
    const double fill = static_cast<double>(qf) * in.invQ;
    const double tb = L.tb, bb = L.bb;
    const double o = (node + tb + bb) / in.binArea - static_cast<double>(in.g.target);
    *over += o > 0.0 ? o * in.binArea : 0.0;
    if (in.dbgNode) in.dbgNode[b] = qn;
    if (in.dbgFill) in.dbgFill[b] = qf;
    // a constant shift only changes the DC term, which the solve discards;
    // it keeps the transform inputs small (smaller FP32 rounding)
    return static_cast<T>((node + fill + tb + bb) / in.binArea - in.rhoShift);
}

template <typename T>
__device__ inline T finalizeBin(const DensityIn<T>& in, int b, double* over) {
    BinLoads L;
    finalizeLoad<T>(in, b, L);
    return finalizeFinish<T>(in, b, L, over);
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
