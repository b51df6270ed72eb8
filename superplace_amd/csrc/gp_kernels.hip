// Per-evaluation kernels of the ePlace iteration for gfx950.
//
// Determinism: the density scatter accumulates 64-bit fixed-point integers
// with atomics (integer addition is associative), and every FP64 reduction is
// a fixed-shape tree: a shuffle tree per block, one partial per block written
// with a plain store, and k_reduce summing the partials in block order.  No
// reduction uses atomics or fences, so results are bitwise reproducible and
// no block has to publish its writes across the chip's L2 caches mid-kernel.
#include <cmath>
#include <stdexcept>

#include "batch.h"
#include "gp_kernels.h"
#include "wa_device.h"
#include "density_device.h"

namespace sphip {
namespace {

constexpr int kReduceBatch = 8;   // block partials per thread loaded together by k_reduce

// ------------------------------------------------------------------ footprint
template <typename T>
struct Footprint {
    int bx0, bx1, by0, by1;
    T lx, ly, hx, hy, wt;
};

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// BinGrid::footprintOf + chargeWeight (bin_grid.cpp:130-170): cells narrower
// (shorter) than a bin are widened to one bin, and the charge is scaled by
// w/stepX (h/stepY); macros are additionally scaled by the target density.
// Contraction is off so the rounding matches the reference's C++ exactly.
template <typename T>
__device__ inline Footprint<T> footprintOf(T cx, T cy, T w, T h, uint8_t f, const GridGeom<T>& g) {
#pragma clang fp contract(off)
    Footprint<T> fp;
    T lx = cx - T(0.5) * w, hx = cx + T(0.5) * w;
    T ly = cy - T(0.5) * h, hy = cy + T(0.5) * h;
    T sx = T(1), sy = T(1);
    if (w < g.stepX) {
        sx = w / g.stepX;
        lx = cx - T(0.5) * g.stepX;
        hx = cx + T(0.5) * g.stepX;
    }
    if (h < g.stepY) {
        sy = h / g.stepY;
        ly = cy - T(0.5) * g.stepY;
        hy = cy + T(0.5) * g.stepY;
    }
    const T scale = sx * sy;
    fp.wt = (f & F_FILLER) ? scale : ((f & F_MACRO) ? scale * g.target : scale);
    fp.lx = lx; fp.hx = hx; fp.ly = ly; fp.hy = hy;
    fp.bx0 = clampi(static_cast<int>((lx - g.originX) / g.stepX), 0, g.dim - 1);
    fp.bx1 = clampi(static_cast<int>((hx - g.originX) / g.stepX), 0, g.dim - 1);
    fp.by0 = clampi(static_cast<int>((ly - g.originY) / g.stepY), 0, g.dim - 1);
    fp.by1 = clampi(static_cast<int>((hy - g.originY) / g.stepY), 0, g.dim - 1);
    return fp;
}

// overlapArea(bin, rect) * w (geometry.h:35-39, bin_grid.cpp:328-334).
template <typename T>
__device__ inline T binCharge(int bx, int by, const Footprint<T>& fp, const GridGeom<T>& g) {
#pragma clang fp contract(off)
    const T blx = g.originX + static_cast<T>(bx) * g.stepX;
    const T bly = g.originY + static_cast<T>(by) * g.stepY;
    const T dx = fmin(blx + g.stepX, fp.hx) - fmax(blx, fp.lx);
    const T dy = fmin(bly + g.stepY, fp.hy) - fmax(bly, fp.ly);
    const T ov = (dx > T(0) && dy > T(0)) ? dx * dy : T(0);
    return ov * fp.wt;
}

__device__ inline long long toFixed(float v, float s) { return __float2ll_rn(v * s); }
__device__ inline long long toFixed(double v, double s) { return __double2ll_rn(v * s); }

// ------------------------------------------------------------------ spatial order
// Morton interleave of 6-bit tile coordinates (dim <= 1024 -> 64 x 64 tiles).
__device__ inline unsigned mortonKey(unsigned tx, unsigned ty) {
    unsigned k = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) k |= ((tx >> b) & 1u) << (2 * b) | ((ty >> b) & 1u) << (2 * b + 1);
    return k;
}

template <typename T>
__device__ inline unsigned tileKeyOf(T x, T y, const GridGeom<T>& g) {
    const int bx = clampi(static_cast<int>((x - g.originX) / g.stepX), 0, g.dim - 1);
    const int by = clampi(static_cast<int>((y - g.originY) / g.stepY), 0, g.dim - 1);
    return mortonKey(static_cast<unsigned>(bx / kTileBins), static_cast<unsigned>(by / kTileBins));
}

__device__ inline unsigned mortonCoord(unsigned k) {   // the even bits of k
    unsigned c = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) c |= ((k >> (2 * b)) & 1u) << b;
    return c;
}

// Counting sort by tile key.  Each block takes kBlock nodes of the previous
// order (nearly sorted already, so a block sees few keys), counts its keys in
// an LDS histogram (remembering the distinct keys), then touches each
// distinct key once in global memory: pass 1 adds the block's counts, pass 2
// reserves the block's range at the key's cursor and hands out ranks from
// LDS.  The order within a tile is arbitrary, which the integer density sums
// do not see.
struct TileHist {
    int count[kTileKeys];
    int keys[kBlock];
    int nKeys;
};

template <typename T>
__device__ inline unsigned blockTileHist(const DevDesign<T>& d, const GridGeom<T>& g, TileHist& h,
                                         int chunk, int* outRec, typename Vec2<T>::type* outXY) {
    for (int k = threadIdx.x; k < d.tiles.nKeys; k += kBlock) h.count[k] = 0;
    if (threadIdx.x == 0) h.nKeys = 0;
    __syncthreads();
    const int t = chunk * kBlock + threadIdx.x;
    unsigned key = 0xFFFFFFFFu;
    if (t < d.nDens) {
        const int rec = d.recIdA[t];
        const auto p = load2(d.xy, rec & kRecNode);
        *outRec = rec;
        *outXY = p;
        key = tileKeyOf<T>(p.x, p.y, g);
        if (atomicAdd(&h.count[key], 1) == 0) h.keys[atomicAdd(&h.nKeys, 1)] = key;
    }
    __syncthreads();
    return key;
}

// Exclusive prefix sums of src[0..n) into dst (LDS, kTileKeys entries),
// n <= kTileKeys; src may alias dst.  Returns the total.  Block-wide.
__device__ inline int blockExclusiveScan(const int* src, int n, int* dst, int* waveTot) {
    const int per = (n + kBlock - 1) / kBlock;
    const int t0 = threadIdx.x * per;
    int sum = 0;
    for (int j = 0; j < per; ++j) {
        const int k = t0 + j;
        const int v = k < n ? src[k] : 0;
        dst[k] = v;
        sum += v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    if (lane == 63) waveTot[wave] = inc;
    __syncthreads();
    int run = inc - sum, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) run += waveTot[w];
        total += waveTot[w];
    }
    for (int j = 0; j < per; ++j) {
        const int k = t0 + j;
        const int v = dst[k];
        dst[k] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

// Both passes stride over chunks of kBlock records with a bounded grid, so
// that a skipped sort (ctl->sortNext clear) costs a small launch.
// Block `vb` of `nvb` (the kernels below take their block index explicitly,
// so that a batched launch can run the blocks of several instances).
template <typename T>
__device__ inline void tileCountBody(const DevDesign<T>& d, const GridGeom<T>& g, const Ctl* ctl,
                                     int vb, int nvb) {
    if (ctl && (ctl->done || !ctl->sortNext)) return;
    __shared__ TileHist h;
    const int chunks = (d.nDens + kBlock - 1) / kBlock;
    for (int chunk = vb; chunk < chunks; chunk += nvb) {
        int rec = 0;
        typename Vec2<T>::type p;
        blockTileHist<T>(d, g, h, chunk, &rec, &p);
        for (int q = threadIdx.x; q < h.nKeys; q += kBlock) {
            const int key = h.keys[q];
            atomicAdd(&d.tiles.count[key], h.count[key]);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_tile_count(DevDesign<T> d, GridGeom<T> g,
                                                       const Ctl* ctl) {
    tileCountBody<T>(d, g, ctl, blockIdx.x, gridDim.x);
}

// Pass 2: every block scans the counts itself (a few KB from L2), so no
// separate scan launch is needed.  The grid has one block more than the
// counting pass: that last block publishes the key starts and the extra
// scatter blocks of crowded tiles, and places no records.
template <typename T>
__device__ inline void tilePlaceBody(const DevDesign<T>& d, const GridGeom<T>& g, const Ctl* ctl,
                                     int vb, int nvb) {
    if (ctl && (ctl->done || !ctl->sortNext)) return;
    __shared__ TileHist h;
    __shared__ int prefix[kTileKeys];
    __shared__ int waveTot[kBlock / 64];
    const TileGrid& tg = d.tiles;
    blockExclusiveScan(tg.count, tg.nKeys, prefix, waveTot);
    // the last block publishes the key starts and the extra scatter blocks
    // of crowded tiles while the others place the records
    const int placers = nvb - 1;
    if (vb == placers) {
        for (int k = threadIdx.x; k < tg.nKeys; k += kBlock) {
            tg.start[k] = prefix[k];
            tg.size[k] = tg.count[k];
        }
        __syncthreads();
        for (int k = threadIdx.x; k < tg.nKeys; k += kBlock) {
            const int c = tg.count[k];
            prefix[k] = c > tg.chunk ? (c - 1) / tg.chunk : 0;
        }
        __syncthreads();
        const int total = blockExclusiveScan(prefix, tg.nKeys, prefix, waveTot);
        for (int k = threadIdx.x; k < tg.nKeys; k += kBlock) tg.extraStart[k] = prefix[k];
        if (threadIdx.x == 0) tg.extraStart[tg.nKeys] = total;
        return;
    }
    const int chunks = (d.nDens + kBlock - 1) / kBlock;
    for (int chunk = vb; chunk < chunks; chunk += placers) {
        int rec = 0;
        typename Vec2<T>::type p;
        const unsigned key = blockTileHist<T>(d, g, h, chunk, &rec, &p);
        for (int q = threadIdx.x; q < h.nKeys; q += kBlock) {
            const int k = h.keys[q];
            h.count[k] = prefix[k] + atomicAdd(&tg.fill[k], h.count[k]);
        }
        __syncthreads();
        if (key != 0xFFFFFFFFu) {
            const int t = chunk * kBlock + threadIdx.x;
            const int pos = atomicAdd(&h.count[key], 1);
            const auto wh = load2(d.recWHA, t);
            d.recIdB[pos] = rec;
            store2(d.recWHB, pos, wh.x, wh.y);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_tile_place(DevDesign<T> d, GridGeom<T> g,
                                                       const Ctl* ctl) {
    tilePlaceBody<T>(d, g, ctl, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------ (a) scatter
// Blocks [0, tiles) own one tile each and take its first CHUNK nodes
// (CHUNK = the grid's TileGrid::chunk, which the sort's extra-block prefix
// also uses);
// blocks past that take chunk j >= 1 of a crowded tile (found in extraStart).
// A block accumulates its nodes' fixed-point charges in the tile's window in
// LDS; an owner stores the whole window to the stage (plain stores, every
// evaluation, zeros included), an extra block adds its non-zero window bins
// to the atomic grids.  Charges outside the window (nodes wider than the
// halo) go to the atomic grids directly.  The block also copies its part of
// the records back (A = B) for the next sort.
template <typename T, int CHUNK>
__device__ inline void scatterBody(const DevDesign<T>& d, const GridGeom<T>& g, T qscale,
                                   unsigned long long* nodeQ, unsigned long long* fillQ,
                                   int sorted, double* partials, const Ctl* ctl, int vb) {
    if (ctl) {
        if (ctl->done) return;
        sorted = sorted && ctl->sortNext;
    }
    static_assert(CHUNK % kBlock == 0, "whole records per thread");
    constexpr int PER = CHUNK / kBlock;
    constexpr int WIN2 = kTileWin * kTileWin;
    __shared__ unsigned long long win[2 * WIN2];   // node window, then filler window
    const TileGrid& tg = d.tiles;
    const int nTiles = tg.tilesX * tg.tilesY;
    int tx, ty, key, chunk;
    if (vb < nTiles) {
        tx = vb / tg.tilesY;
        ty = vb - tx * tg.tilesY;
        key = static_cast<int>(mortonKey(tx, ty));
        chunk = 0;
    } else {
        const int e = vb - nTiles;
        if (e >= tg.extraStart[tg.nKeys]) {   // no chunk for this block
            if (threadIdx.x == 0) partials[vb] = 0.0;
            return;
        }
        int lo = 0, hi = tg.nKeys - 1;   // last key with extraStart <= e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tg.extraStart[mid] <= e) lo = mid; else hi = mid - 1;
        }
        key = lo;
        chunk = e - tg.extraStart[lo] + 1;
        tx = static_cast<int>(mortonCoord(key));
        ty = static_cast<int>(mortonCoord(key >> 1));
    }
    for (int e = threadIdx.x; e < 2 * WIN2; e += kBlock) win[e] = 0;
    __syncthreads();
    const int cnt = tg.size[key], st = tg.start[key];
    const int lo = st + chunk * CHUNK;
    const int hi = st + min(cnt, (chunk + 1) * CHUNK);
    const int wx0 = tx * kTileBins - kTileHalo, wy0 = ty * kTileBins - kTileHalo;
    Footprint<T> fp[PER];
    uint8_t fl[PER];
    bool valid[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int t = lo + threadIdx.x + j * kBlock;
        valid[j] = t < hi;
        if (!valid[j]) continue;
        int rec;
        typename Vec2<T>::type wh, p;
        if (sorted) {
            rec = d.recIdB[t];
            wh = load2(d.recWHB, t);
            p = load2(d.xy, rec & kRecNode);
            d.recIdA[t] = rec;
            store2(d.recWHA, t, wh.x, wh.y);
        } else {
            rec = d.recIdA[t];
            wh = load2(d.recWHA, t);
            p = load2(d.xy, rec & kRecNode);
        }
        fl[j] = static_cast<uint8_t>(((rec & kRecFiller) ? F_FILLER : 0) |
                                     ((rec & kRecMacro) ? F_MACRO : 0));
        fp[j] = footprintOf<T>(p.x, p.y, wh.x, wh.y, fl[j], g);
    }
    double spill[1] = {0.0};   // nodes with charge outside the window
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        if (!valid[j]) continue;
        const bool filler = fl[j] & F_FILLER;
        unsigned long long* grid = filler ? fillQ : nodeQ;
        unsigned long long* w = win + (filler ? WIN2 : 0);
        bool out = false;
        for (int bx = fp[j].bx0; bx <= fp[j].bx1; ++bx)
            for (int by = fp[j].by0; by <= fp[j].by1; ++by) {
                const T ov = binCharge<T>(bx, by, fp[j], g);
                if (ov <= T(0)) continue;
                const unsigned long long q = static_cast<unsigned long long>(toFixed(ov, qscale));
                const int lx = bx - wx0, ly = by - wy0;
                if (lx >= 0 && lx < kTileWin && ly >= 0 && ly < kTileWin) {
                    atomicAdd(&w[lx * kTileWin + ly], q);
                } else {
                    atomicAdd(&grid[static_cast<long>(bx) * g.dim + by], q);
                    out = true;
                }
            }
        if (out) spill[0] += 1.0;
    }
    blockPartial<1>(spill, partials, vb);   // ends with a barrier
    if (chunk == 0) {
        unsigned long long* out = tg.stage + static_cast<long>(vb) * 2 * WIN2;
        for (int e = threadIdx.x; e < 2 * WIN2; e += kBlock) out[e] = win[e];
    } else {
        for (int e = threadIdx.x; e < 2 * WIN2; e += kBlock) {
            const unsigned long long v = win[e];
            if (!v) continue;
            const int r = e < WIN2 ? e : e - WIN2;
            const long gi = static_cast<long>(wx0 + r / kTileWin) * g.dim + wy0 + r % kTileWin;
            atomicAdd(&(e < WIN2 ? nodeQ : fillQ)[gi], v);
        }
    }
}

template <typename T, int CHUNK>
__global__ __launch_bounds__(kBlock) void k_scatter(DevDesign<T> d, GridGeom<T> g, T qscale,
                                                    unsigned long long* nodeQ,
                                                    unsigned long long* fillQ, int sorted,
                                                    double* partials, const Ctl* ctl) {
    scatterBody<T, CHUNK>(d, g, qscale, nodeQ, fillQ, sorted, partials, ctl, blockIdx.x);
}

// scatterBody for the grid's chunk (one of scatterChunkOk's).
template <typename T>
__device__ inline void scatterAnyChunk(const DevDesign<T>& d, const GridGeom<T>& g, T qscale,
                                       unsigned long long* nodeQ, unsigned long long* fillQ,
                                       int sorted, double* partials, const Ctl* ctl, int vb) {
    switch (d.tiles.chunk) {
        case 256: scatterBody<T, 256>(d, g, qscale, nodeQ, fillQ, sorted, partials, ctl, vb); break;
        case 1024: scatterBody<T, 1024>(d, g, qscale, nodeQ, fillQ, sorted, partials, ctl, vb); break;
        case 2048: scatterBody<T, 2048>(d, g, qscale, nodeQ, fillQ, sorted, partials, ctl, vb); break;
        default: scatterBody<T, 512>(d, g, qscale, nodeQ, fillQ, sorted, partials, ctl, vb); break;
    }
}

// ------------------------------------------------------------------ rho + tau
// finalizeBin over every bin (density_device.h), rho into the padded layout,
// one tau partial per block; clears the tile counters.
template <typename T>
__device__ inline void finalizeBody(const DensityIn<T>& in, T* rho, const Ctl* ctl, int vb) {
    if (ctl && ctl->done) return;
    const int b = vb * kBlock + threadIdx.x;
    const int nb = in.g.dim * in.g.dim;
    double v[1] = {0.0};
    clearTileKeys(in.tg, b);
    if (b < nb) {
        const T r = finalizeBin<T>(in, b, &v[0]);
        const int bx = b / in.g.dim, by = b - bx * in.g.dim;
        rho[static_cast<long>(bx) * in.g.ld + by] = r;
    }
    blockPartial<1>(v, in.partials, vb);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_finalize(DensityIn<T> in, T* rho, const Ctl* ctl) {
    finalizeBody<T>(in, rho, ctl, blockIdx.x);
}

// WaWirelength::compute (wa_wirelength.cpp:119-225) as one pin-parallel pass
// over the net-ordered slots: per net the extrema, max-shifted exponentials,
// centred sums and WA value (degree 2..ignoreDeg), the HPWL of every net of
// degree >= 2 (hpwl.cpp:20-37), and each pin's gradient into pinGrad at the
// pin's node-pin entry.  Each wavefront belongs to one size class.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_wa_pins(WaTask<T> w) {
    if (w.ctl && w.ctl->done) return;
    waBlockTask<T>(w, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------ (c)+(d) node pass
// For each optimised node k (node i = optIdx[k]):
//  * WA gradient: the sum of its pins' gradients, in node-pin order
//    (wa_wirelength.cpp:183-225),
//  * density force F_i = sum_bins q * xi (bin_grid.cpp:395-426),
//  * preconditioned gradient grad = (dWL - lambda F) / max(1, deg + lambda area)
//    (eplace.cpp:162-185),
//  * FP64 partial sums: sum|dWL|, sum|F| (eplace.cpp:168-190) and the
//    Nesterov step statistics ||u-up||^2, ||g-gp||^2, (u-up).(g-gp), ||g||^2
//    (nesterov.cpp:25-44), plus a count of non-finite gradients.
// Node index of optimised node k (optIdx, computed when it is affine).
template <typename T>
__device__ inline int optNode(const DevDesign<T>& d, int k) {
    if (d.optAffine) return k < d.optSplit ? k : k + d.optShift;
    return d.optIdx[k];
}

// Block b of a node-pass or update grid of nblocks blocks (nodeBlocks):
// its first opt index, the end of its part, and its logical block (its
// partial's slot).  Cells take blocks [0, cellBlocks), fillers the rest;
// xcdBlock orders each part.
struct NodeRange {
    int k0, end, slot;
};
template <typename T>
__device__ inline NodeRange nodeRange(const DevDesign<T>& d, int b, int nblocks) {
    NodeRange r;
    if (b < d.cellBlocks) {
        r.slot = xcdBlock(b, d.cellBlocks);
        r.k0 = r.slot * kBlock;
        r.end = d.nCells;
    } else {
        const int f = xcdBlock(b - d.cellBlocks, nblocks - d.cellBlocks);
        r.slot = d.cellBlocks + f;
        r.k0 = d.nCells + f * kBlock;
        r.end = d.nOpt;
    }
    return r;
}

// Loads are ordered by dependence: the node's own data and the Nesterov
// vectors first (one memory round trip), then its footprint's field bins and
// its pin entries (a second; pin entries eight at a time).  The sums run in
// the same order as a plain sequential loop over pins and bins.
template <typename T>
__device__ inline void nodeGradBody(const DevDesign<T>& d, const GridGeom<T>& g, const T* exy,
                                    T invG, float lambda, const T* u, const T* up, const T* gp,
                                    T* grad, T* dbgForce, T* dbgWl, double* partials,
                                    const Ctl* ctl, const BufTable<T>& bufs, int b,
                                    int nblocks) {
    if (ctl) {
        if (ctl->done) return;
        lambda = ctl->lambda;
        invG = invGammaOf<T>(ctl->gamma);
        u = pickBuf(bufs, ctl->nxt[B_U]);
        up = pickBuf(bufs, ctl->nxt[B_UP]);
        gp = pickBuf(bufs, ctl->nxt[B_GP]);
        grad = pickBuf(bufs, ctl->nxt[B_G]);
    }
    const NodeRange nr = nodeRange(d, b, nblocks);
    const int k = nr.k0 + threadIdx.x;
    const int n = d.nOpt;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (k < nr.end) {
        const int i = optNode(d, k);
        const double u0 = u[k], u1 = u[n + k];
        const double up0 = up ? static_cast<double>(up[k]) : 0.0;
        const double up1 = up ? static_cast<double>(up[n + k]) : 0.0;
        const double gp0 = gp ? static_cast<double>(gp[k]) : 0.0;
        const double gp1 = gp ? static_cast<double>(gp[n + k]) : 0.0;
        const auto pi = load2(d.xy, i);
        const T wi = d.w[i], hi = d.h[i];
        const uint8_t f = d.flags[i];
        const bool hasWl = i < d.numMov;
        const int t0 = hasWl ? d.nodeStart[i] : 0, t1 = hasWl ? d.nodeStart[i + 1] : 0;
        const T xi = pi.x, yi = pi.y;
        // the footprint's first four bins (x-major order), loads issued now
        const bool dens = !(f & (F_FIXED | F_NI));
        Footprint<T> fp{};
        int nby = 1, nbin = 0;
        if (dens) {
            fp = footprintOf<T>(xi, yi, wi, hi, f, g);
            nby = fp.by1 - fp.by0 + 1;
            nbin = (fp.bx1 - fp.bx0 + 1) * nby;
        }
        T q[4], fxv[4], fyv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            q[j] = T(0);
            if (j < nbin) {
                const int bx = fp.bx0 + j / nby, by = fp.by0 + j % nby;
                q[j] = binCharge<T>(bx, by, fp, g);
                if (q[j] > T(0)) {
                    const auto fe = load2(exy, static_cast<long>(bx) * g.ld + by);
                    fxv[j] = fe.x;
                    fyv[j] = fe.y;
                }
            }
        }
        // WA gradient: the node's pin entries, eight loads in flight at a time
        T gx = 0, gy = 0;
        for (int t = t0; t < t1; t += 8) {
            typename Vec2<T>::type pg[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (t + j < t1) pg[j] = load2(d.pinGrad, t + j);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (t + j < t1) {
                    gx += pg[j].x;
                    gy += pg[j].y;
                }
        }
        // density force over the footprint
        T fx = 0, fy = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (q[j] > T(0)) {
                fx += q[j] * fxv[j];
                fy += q[j] * fyv[j];
            }
        for (int e0 = 4; e0 < nbin; e0 += 4) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                q[j] = T(0);
                if (e0 + j < nbin) {
                    const int bx = fp.bx0 + (e0 + j) / nby, by = fp.by0 + (e0 + j) % nby;
                    q[j] = binCharge<T>(bx, by, fp, g);
                    if (q[j] > T(0)) {
                        const auto fe = load2(exy, static_cast<long>(bx) * g.ld + by);
                        fxv[j] = fe.x;
                        fyv[j] = fe.y;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (q[j] > T(0)) {
                    fx += q[j] * fxv[j];
                    fy += q[j] * fyv[j];
                }
        }
        const T lam = static_cast<T>(lambda);
        // the preconditioner's pin count and area (eplace.cpp:178): the
        // node's pin range and w h (exact in FP64)
        const double pdeg = static_cast<double>(static_cast<float>(t1 - t0));
        const double parea = static_cast<double>(wi) * static_cast<double>(hi);
        const T precond = static_cast<T>(1.0 / fmax(1.0, pdeg + static_cast<double>(lambda) * parea));
        const T g0 = precond * (gx - lam * fx);
        const T g1 = precond * (gy - lam * fy);
        grad[k] = g0;
        grad[n + k] = g1;
        if (dbgForce) { dbgForce[k] = fx; dbgForce[n + k] = fy; }
        if (dbgWl) { dbgWl[k] = gx; dbgWl[n + k] = gy; }
        v[0] = hasWl ? fabs(static_cast<double>(gx)) + fabs(static_cast<double>(gy)) : 0.0;
        v[1] = fabs(static_cast<double>(fx)) + fabs(static_cast<double>(fy));
        const double du0 = u0 - up0, du1 = u1 - up1;
        const double dg0 = static_cast<double>(g0) - gp0;
        const double dg1 = static_cast<double>(g1) - gp1;
        v[2] = du0 * du0 + du1 * du1;
        v[3] = dg0 * dg0 + dg1 * dg1;
        v[4] = du0 * dg0 + du1 * dg1;
        v[5] = static_cast<double>(g0) * g0 + static_cast<double>(g1) * g1;
        v[6] = (isfinite(g0) ? 0.0 : 1.0) + (isfinite(g1) ? 0.0 : 1.0);
    }
    blockPartial<7>(v, partials, nr.slot);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_node_grad(DevDesign<T> d, GridGeom<T> g, const T* exy,
                                                      T invG, float lambda,
                                                      const T* u, const T* up, const T* gp, T* grad,
                                                      T* dbgForce, T* dbgWl, double* partials,
                                                      const Ctl* ctl, BufTable<T> bufs) {
    nodeGradBody<T>(d, g, exy, invG, lambda, u, up, gp, grad, dbgForce, dbgWl, partials, ctl,
                    bufs, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------ (d) decisions
// computeGamma (wa_wirelength.cpp:18-29) and lambdaMultiplier (lambda_schedule.h:22-25).
// The FP32 arithmetic is not contracted (the reference's is not), and the
// power is 10^e in FP64 rounded once to FP32: the correctly rounded value,
// which the host loop (pow10f, engine.hip) takes too.
__device__ inline float gammaOf(float tau, float binStep) {
#pragma clang fp contract(off)
    float gamma = 8.0f * binStep;
    if (tau > 1.0f) gamma *= 10.0f;
    else if (tau < 0.1f) gamma *= 0.1f;
    else {
        const float e = (20.0f / 9.0f) * (tau - 0.1f) - 1.0f;
        gamma *= static_cast<float>(pow(10.0, static_cast<double>(e)));
    }
    return fmaxf(gamma, 1e-4f);
}

__device__ inline float lambdaMul(double dH, double ref) {
#pragma clang fp contract(off)
    const double mu = pow(1.1, -dH / ref + 1.0);
    return static_cast<float>(fmin(1.05, fmax(0.95, mu)));
}

// One thread, once the scalars of an evaluation are summed (sc = EvalScalars):
// the backtracking check (nesterov.cpp:146-169), the step size (:88,
// :221-240), the momentum (:92-93), then - for a completed step - the
// schedules (eplace.cpp:371-380), the metrics row (:384-399), the stop test
// (:411-419), the tau-ladder snapshots and stagnation rollback (:421-473) and
// the iteration limit.  The update kernel executes the chosen mode with the
// `cur` indices; `nxt` holds the indices after the step's rotation
// (nesterov.cpp:109-118).
// The loop state's scalars and buffer roles in registers for the decisions
// (thread 0 of the reduction otherwise makes a chain of dependent LDS round
// trips through Ctl); the tau-ladder arrays, indexed by counts, stay in Ctl.
#define SPHIP_CTL_SCALARS(X)                                                              \
    X(mode) X(done) X(status) X(iter) X(tries) X(maxIter) X(maxBt) X(backtrack) X(useBB)   \
    X(stage) X(window) X(evals) X(finalBuf) X(rollbackIter) X(snapWrite) X(snapSrc)        \
    X(nSnap) X(nFree) X(a) X(alpha) X(alpha0) X(m) X(momPrev) X(lambda) X(gamma)           \
    X(targetTau) X(stepX) X(minGain) X(bestTau) X(refTau) X(sinceImprove) X(havePrev)      \
    X(sortNext) X(spillLimit) X(prevHpwl) X(effRef)
struct CtlRegs {
    int cur[B_ROLES], nxt[B_ROLES];
#define SPHIP_DECL(f) decltype(Ctl::f) f;
    SPHIP_CTL_SCALARS(SPHIP_DECL)
#undef SPHIP_DECL
};

template <typename T>
__device__ __attribute__((always_inline)) inline void decideRegs(CtlRegs& st, Ctl& c,
                                                                 const double* sc, DevRow* rows) {
#pragma clang fp contract(off)
#pragma unroll
    for (int r = 0; r < B_ROLES; ++r) st.cur[r] = st.nxt[r];
    st.snapWrite = -1;
    const double tau = sc[0], wl = sc[1], hpwl = sc[2], sumWl = sc[3], sumDen = sc[4];
    const double dx = sqrt(sc[5]), dg = sqrt(sc[6]), sy = sc[7], nonfin = sc[9];
    if (!isfinite(tau) || !isfinite(hpwl) || !isfinite(wl) || !isfinite(sumWl) ||
        !isfinite(sumDen) || nonfin > 0.0) {
        st.done = 1;
        st.status = kNonFinite;
        st.mode = kModeNone;
        return;
    }
    ++st.evals;
    if (st.backtrack && st.iter > 0 && st.tries < st.maxBt - 1 && dx > 0.0 && dg > 0.0 &&
        isfinite(dg)) {
        const double lip = dx / dg;
        if (!(lip > 0.95 * st.alpha)) {
            st.alpha = static_cast<double>(static_cast<T>(lip));
            st.m = st.momPrev;
            st.mode = kModeRedo;   // v, u rewritten in place: nxt = cur
            ++st.tries;
            return;
        }
    }
    st.tries = 0;
    if (st.iter == 0) {
        st.alpha = st.alpha0;
    } else if (dx > 0.0 && dg > 0.0) {
        double next = dx / dg;
        if (st.useBB) {
            const double yy = dg * dg;
            if (yy > 0.0) {
                const double bb = sy / yy;
                if (bb > 0.0 && isfinite(bb)) next = bb;
            }
        }
        st.alpha = static_cast<double>(static_cast<T>(next));
    }
    const T aT = static_cast<T>(st.a);
    const T aNext = T(0.5) * (T(1) + sqrt(T(4) * aT * aT + T(1)));
    const T mom = (aT - T(1)) / aNext;
    st.m = static_cast<double>(mom);
    st.momPrev = st.m;
    st.a = static_cast<double>(aNext);
    st.mode = kModeStep;
    st.nxt[B_U] = st.cur[B_UN];
    st.nxt[B_V] = st.cur[B_VN];
    st.nxt[B_UP] = st.cur[B_U];
    st.nxt[B_VP] = st.cur[B_V];
    st.nxt[B_UN] = st.cur[B_UP];
    st.nxt[B_VN] = st.cur[B_VP];
    st.nxt[B_G] = st.cur[B_GP];
    st.nxt[B_GP] = st.cur[B_G];
    const int lastEval = st.cur[B_U];   // the point this evaluation used

    const float evalLambda = st.lambda, evalGamma = st.gamma;
    const float ftau = static_cast<float>(tau);
    const double dH = st.havePrev ? hpwl - st.prevHpwl : 0.0;
    const float mu = st.havePrev ? lambdaMul(dH, st.effRef) : 1.f;
    if (st.havePrev) st.lambda *= mu;
    st.prevHpwl = hpwl;
    st.havePrev = 1;
    st.gamma = gammaOf(ftau, st.stepX);
    if (!isfinite(st.lambda) || !isfinite(st.gamma)) {
        st.done = 1;
        st.status = kNonFinite;
        st.mode = kModeNone;
        return;
    }
    const int it = st.iter;
    if (rows) {
        DevRow& r = rows[it];
        r.iter = it;
        r.evaluations = static_cast<int>(st.evals);
        r.hpwl = hpwl;
        r.overflow = ftau;
        r.lambda = st.lambda;
        r.gamma = st.gamma;
        r.step = st.alpha;
        r.gradWl = static_cast<float>(sumWl);
        r.gradDen = static_cast<float>(sumDen);
        r.evalLambda = evalLambda;
        r.evalGamma = evalGamma;
        r.dH = dH;
        r.mu = mu;
        r.ticks = __builtin_amdgcn_s_memrealtime();
    }
    st.iter = it + 1;
    if (st.stage != 1 && ftau < st.targetTau) {   // deliver the look-ahead point
        st.done = 1;
        st.status = kConverged;
        st.finalBuf = lastEval;
        st.mode = kModeNone;
        return;
    }
    if (st.stage != 1 && st.window > 0) {
        if (ftau < st.refTau * (1.f - st.minGain)) {
            st.refTau = ftau;
            st.sinceImprove = 0;
        } else {
            ++st.sinceImprove;
        }
        st.bestTau = fminf(st.bestTau, ftau);
        const float tol = st.bestTau * (1.f + st.minGain);
        auto dropFront = [&]() {
            c.freeSnap[st.nFree++] = c.snapBuf[0];
            for (int s = 1; s < st.nSnap; ++s) {
                c.snapBuf[s - 1] = c.snapBuf[s];
                c.snapTau[s - 1] = c.snapTau[s];
                c.snapHpwl[s - 1] = c.snapHpwl[s];
                c.snapIter[s - 1] = c.snapIter[s];
            }
            --st.nSnap;
        };
        while (st.nSnap > 0 && c.snapTau[0] > tol) dropFront();
        if (st.nSnap == 0 || ftau * (1.f + 0.5f * st.minGain) <= c.snapTau[st.nSnap - 1]) {
            if (st.nSnap == kMaxSnaps) dropFront();
            const int b = c.freeSnap[--st.nFree];
            c.snapBuf[st.nSnap] = b;
            c.snapTau[st.nSnap] = ftau;
            c.snapHpwl[st.nSnap] = hpwl;
            c.snapIter[st.nSnap] = it;
            ++st.nSnap;
            st.snapWrite = b;
            st.snapSrc = lastEval;
        }
        if (st.sinceImprove >= st.window) {
            int pick = 0;
            for (int s = 1; s < st.nSnap; ++s)
                if (c.snapHpwl[s] < c.snapHpwl[pick]) pick = s;
            st.done = 1;
            st.status = kStagnated;
            st.finalBuf = c.snapBuf[pick];
            st.rollbackIter = c.snapIter[pick];
            // the update still runs this slot so that a snapshot taken now is copied
            return;
        }
    }
    if (st.iter >= st.maxIter) {
        st.done = 1;
        st.status = kIterLimit;
        st.finalBuf = st.cur[B_VN];   // v after this step (pos = v, nesterov.cpp:121)
    }
}

template <typename T>
__device__ void decideStep(Ctl& c, const double* sc, DevRow* rows) {
    CtlRegs st;
#pragma unroll
    for (int r = 0; r < B_ROLES; ++r) {
        st.cur[r] = c.cur[r];
        st.nxt[r] = c.nxt[r];
    }
#define SPHIP_LOAD(f) st.f = c.f;
    SPHIP_CTL_SCALARS(SPHIP_LOAD)
#undef SPHIP_LOAD
    decideRegs<T>(st, c, sc, rows);
#pragma unroll
    for (int r = 0; r < B_ROLES; ++r) {
        c.cur[r] = st.cur[r];
        c.nxt[r] = st.nxt[r];
    }
#define SPHIP_STORE(f) c.f = st.f;
    SPHIP_CTL_SCALARS(SPHIP_STORE)
#undef SPHIP_STORE
}

// Single block: sums the regions this evaluation wrote into EvalScalars
// [tau, wl, hpwl, sum_wl, sum_den, dx2, dg2, sy, gg, nonfinite], then (device
// loop) takes the decisions.
template <typename T>
__device__ inline void reduceBody(const ReduceWs& ws, double denom, int withNode, Ctl* ctl,
                                  DevRow* rows) {
    if (ctl && ctl->done) return;
    // the decision thread works on an LDS copy of the loop state: one
    // cooperative copy in and out instead of dependent global round trips
    constexpr int kCtlWords = (sizeof(Ctl) + 3) / 4;
    __shared__ __attribute__((aligned(16))) unsigned ctlWords[kCtlWords];
    if (ctl)
        for (int w = threadIdx.x; w < kCtlWords; w += blockDim.x)
            ctlWords[w] = reinterpret_cast<const unsigned*>(ctl)[w];
    // all regions in one pass (every load independent), one block sum
    double r[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) r[j] = 0.0;
    {
        const double* p0 = ws.region(0);
        const double* p1 = ws.region(1);
        const double* p2 = ws.region(2);
        const double* p3 = ws.region(3);
        const int n3 = withNode ? ws.blocks[3] : 0;
        const int nmax = max(max(ws.blocks[0], ws.blocks[1]), max(ws.blocks[2], n3));
        // rows of kReduceBatch strides at a time, so that all their loads are
        // in flight together (one memory latency per batch, not per row)
        for (int b0 = threadIdx.x; b0 < nmax; b0 += kReduceBatch * kBlock) {
            double v[kReduceBatch][12];
#pragma unroll
            for (int q = 0; q < kReduceBatch; ++q) {
                const int b = b0 + q * kBlock;
#pragma unroll
                for (int j = 0; j < 12; ++j) v[q][j] = 0.0;
                if (b < ws.blocks[0]) v[q][0] = p0[b];
                if (b < ws.blocks[1]) { v[q][1] = p1[2L * b]; v[q][2] = p1[2L * b + 1]; }
                if (b < ws.blocks[2]) v[q][3] = p2[b];
                if (b < n3)
#pragma unroll
                    for (int j = 0; j < 7; ++j) v[q][5 + j] = p3[7L * b + j];
            }
#pragma unroll
            for (int q = 0; q < kReduceBatch; ++q)
#pragma unroll
                for (int j = 0; j < 12; ++j) r[j] += v[q][j];
        }
    }
    blockSum<12>(r);   // ends with a barrier
    __shared__ double sc[kNumScalars];
    if (threadIdx.x == 0) {
        sc[0] = denom > 0.0 ? r[0] / denom : 0.0;
        sc[1] = r[1];
        sc[2] = r[2];
        for (int j = 0; j < 7; ++j) sc[3 + j] = withNode ? r[5 + j] : 0.0;
        for (int j = 0; j < kNumScalars; ++j) ws.scalars[j] = sc[j];
        if (ctl && withNode) {
            Ctl& c = *reinterpret_cast<Ctl*>(ctlWords);
            c.sortNext = r[3] > c.spillLimit ? 1 : 0;   // re-sort before the next scatter
            decideStep<T>(c, sc, rows);
        }
    }
    __syncthreads();
    if (ctl && withNode)
        for (int w = threadIdx.x; w < kCtlWords; w += blockDim.x)
            reinterpret_cast<unsigned*>(ctl)[w] = ctlWords[w];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_reduce(ReduceWs ws, double denom, int withNode,
                                                   Ctl* ctl, DevRow* rows) {
    reduceBody<T>(ws, denom, withNode, ctl, rows);
}

// ------------------------------------------------------------------ (d) update
// NesterovOptimizer::step / redoIfOvershot arithmetic (nesterov.cpp:97-107,
// :156-165) with EPlace's clamp (eplace.cpp:117-133) and scatter (:97-105).
template <typename T>
__device__ inline void updateBody(const DevDesign<T>& d, const T* A, const T* G, const T* B,
                                  T alpha, T m, T* vOut, T* uOut, const Ctl* ctl,
                                  const BufTable<T>& bufs, int b, int nblocks) {
    const NodeRange nr = nodeRange(d, b, nblocks);
    const int k = nr.k0 + threadIdx.x;
    const int n = d.nOpt;
    if (ctl) {
        // After `done` the mode and indices stay fixed, so a repeated update
        // rewrites the same values (reads never alias writes).
        const int mode = ctl->mode;
        if (mode == kModeNone) return;
        if (k < nr.end && ctl->snapWrite >= 0) {   // tau-ladder snapshot of the evaluated point
            const T* src = pickBuf(bufs, ctl->snapSrc);
            T* dst = pickBuf(bufs, ctl->snapWrite);
            dst[k] = src[k];
            dst[n + k] = src[n + k];
        }
        alpha = static_cast<T>(ctl->alpha);
        m = static_cast<T>(ctl->m);
        const int* c = ctl->cur;
        if (mode == kModeStep) {
            A = pickBuf(bufs, c[B_U]); G = pickBuf(bufs, c[B_G]); B = pickBuf(bufs, c[B_V]);
            vOut = pickBuf(bufs, c[B_VN]); uOut = pickBuf(bufs, c[B_UN]);
        } else {   // redo the previous step from u_prev (nesterov.cpp:156-165)
            A = pickBuf(bufs, c[B_UP]); G = pickBuf(bufs, c[B_GP]); B = pickBuf(bufs, c[B_VP]);
            vOut = pickBuf(bufs, c[B_V]); uOut = pickBuf(bufs, c[B_U]);
        }
    }
    if (k >= nr.end) return;
    const int i = optNode(d, k);
    const T hw = T(0.5) * d.w[i], hh = T(0.5) * d.h[i];
    const T eps = static_cast<T>(1e-3f);
    const T xlo = d.coreLx + hw + eps;
    const T xhi = fmax(xlo, d.coreHx - hw - eps);
    const T ylo = d.coreLy + hh + eps;
    const T yhi = fmax(ylo, d.coreHy - hh - eps);
    T vv[2], uu[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int j = a * n + k;
        const T vn = G ? A[j] - alpha * G[j] : A[j];
        const T un = B ? vn + m * (vn - B[j]) : vn;
        const T lo = a ? ylo : xlo, hi = a ? yhi : xhi;
        vv[a] = fmin(fmax(vn, lo), hi);
        uu[a] = fmin(fmax(un, lo), hi);
        if (vOut) vOut[j] = vv[a];
        if (uOut) uOut[j] = uu[a];
    }
    const T nx = uOut ? uu[0] : vv[0], ny = uOut ? uu[1] : vv[1];
    store2(d.xy, i, nx, ny);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_update(DevDesign<T> d, const T* A, const T* G,
                                                   const T* B, T alpha, T m, T* vOut, T* uOut,
                                                   const Ctl* ctl, BufTable<T> bufs) {
    updateBody<T>(d, A, G, B, alpha, m, vOut, uOut, ctl, bufs, blockIdx.x, gridDim.x);
}


long gridFor(long n) { return (n + kBlock - 1) / kBlock; }

// ------------------------------------------------------------------ batches
template <typename T, int KIND>
__global__ __launch_bounds__(kBlock) void k_slot_batch(const SlotArgs<T>* __restrict__ args,
                                                       const int4* __restrict__ map, int sorted) {
    const int4 m = map[blockIdx.x];
    const SlotArgs<T>& a = args[m.x];
    if constexpr (KIND == BK_SORT_COUNT) {
        tileCountBody<T>(a.d, a.g, a.ctl, m.y, m.z);
    } else if constexpr (KIND == BK_SORT_PLACE) {
        tilePlaceBody<T>(a.d, a.g, a.ctl, m.y, m.z);
    } else if constexpr (KIND == BK_SCATTER) {
        scatterAnyChunk<T>(a.d, a.g, a.qscale, a.nodeQ, a.fillQ, sorted, a.ws.region(2), a.ctl,
                           m.y);
    } else if constexpr (KIND == BK_FINALIZE) {
        finalizeBody<T>(a.din, a.rho, a.ctl, m.y);
    } else if constexpr (KIND == BK_WA) {
        if (a.ctl->done) return;
        waBlockTask<T>(a.wa, m.y, m.z);
    } else if constexpr (KIND == BK_NODE) {
        nodeGradBody<T>(a.d, a.g, a.exy, T(0), 0.f, nullptr, nullptr, nullptr, nullptr, nullptr,
                        nullptr, a.ws.region(3), a.ctl, a.bufs, m.y, m.z);
    } else if constexpr (KIND == BK_REDUCE) {
        reduceBody<T>(a.ws, a.denom, 1, a.ctl, a.rows);
        __syncthreads();
        if (threadIdx.x == 0) a.doneOut[m.x] = a.ctl->done;
    } else if constexpr (KIND == BK_UPDATE) {
        updateBody<T>(a.d, nullptr, nullptr, nullptr, T(0), T(0), nullptr, nullptr, a.ctl, a.bufs,
                      m.y, m.z);
    }
}

}  // namespace

template <typename T>
void launchSlotBatch(int kind, const SlotArgs<T>* args, const int4* map, int blocks, int sorted,
                     hipStream_t s) {
    if (blocks <= 0) return;
    switch (kind) {
        case BK_SORT_COUNT:
            launchKernel(k_slot_batch<T, BK_SORT_COUNT>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_SORT_PLACE:
            launchKernel(k_slot_batch<T, BK_SORT_PLACE>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_SCATTER:
            launchKernel(k_slot_batch<T, BK_SCATTER>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_FINALIZE:
            launchKernel(k_slot_batch<T, BK_FINALIZE>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_WA:
            launchKernel(k_slot_batch<T, BK_WA>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_NODE:
            launchKernel(k_slot_batch<T, BK_NODE>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_REDUCE:
            launchKernel(k_slot_batch<T, BK_REDUCE>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        case BK_UPDATE:
            launchKernel(k_slot_batch<T, BK_UPDATE>, blocks, kBlock, 0, s, args, map, sorted);
            break;
        default:
            throw std::logic_error("launchSlotBatch: unknown kernel");
    }
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
void launchScatter(const DevDesign<T>& d, const GridGeom<T>& g, T qscale,
                   unsigned long long* nodeQ, unsigned long long* fillQ, bool sorted,
                   ReduceWs& ws, const Ctl* ctl, hipStream_t s) {
    const int c = d.tiles.chunk;
    if (!scatterChunkOk(c)) throw std::logic_error("launchScatter: unsupported record chunk");
    const int blocks = d.tiles.tilesX * d.tiles.tilesY + (d.nDens + c - 1) / c;
    ws.blocks[2] = blocks;
    auto kern = c == 256 ? k_scatter<T, 256>
              : c == 1024 ? k_scatter<T, 1024>
              : c == 2048 ? k_scatter<T, 2048> : k_scatter<T, 512>;
    launchKernel(kern, blocks, kBlock, 0, s, d, g, qscale, nodeQ, fillQ, sorted ? 1 : 0,
                 ws.region(2), ctl);
    SPHIP_CHECK(hipGetLastError());
}


template <typename T>
void launchTileSort(const DevDesign<T>& d, const GridGeom<T>& g, const Ctl* ctl, hipStream_t s) {
    if (d.nDens == 0) return;
    if (g.dim > kTileBins * 64) throw std::runtime_error("tile sort: grid wider than 1024 bins");
    const int blocks = static_cast<int>(std::min<long>(gridFor(d.nDens), kSortBlocks));
    launchKernel(k_tile_count<T>, blocks, kBlock, 0, s, d, g, ctl);
    SPHIP_CHECK(hipGetLastError());
    launchKernel(k_tile_place<T>, blocks + 1, kBlock, 0, s, d, g, ctl);   // + the publishing block
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
void launchFinalize(const GridGeom<T>& g, const TileGrid& tg, unsigned long long* nodeQ,
                    unsigned long long* fillQ, const double* terminal, const double* base,
                    double invQ, double binArea, double rhoShift, T* rho, long long* dbgNode,
                    long long* dbgFill, ReduceWs& ws, const Ctl* ctl, hipStream_t s) {
    const long nb = static_cast<long>(g.dim) * g.dim;
    ws.blocks[0] = static_cast<int>(gridFor(std::max<long>(nb, tg.nKeys)));
    const DensityIn<T> in{g, tg, nodeQ, fillQ, terminal, base, invQ, binArea, rhoShift,
                          ws.region(0), dbgNode, dbgFill};
    launchKernel(k_finalize<T>, ws.blocks[0], kBlock, 0, s, in, rho, ctl);
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
int waResidentBlocks() {
    int dev = 0, cus = 0, perCu = 0;
    SPHIP_CHECK(hipGetDevice(&dev));
    SPHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SPHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &perCu, reinterpret_cast<const void*>(k_wa_pins<T>), kBlock, 0));
    return perCu * cus;
}

template <typename T>
void launchWaNets(const DevDesign<T>& d, float gamma, int ignoreDeg, ReduceWs& ws,
                  const Ctl* ctl, hipStream_t s) {
    const T invG = invGammaOf<T>(gamma);
    ws.blocks[1] = d.waBlocks;
    launchKernel(k_wa_pins<T>, ws.blocks[1], kBlock, 0, s, WaTask<T>{d, invG, ignoreDeg, ws.region(1), ctl});
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
void launchNodeGrad(const DevDesign<T>& d, const GridGeom<T>& g, const T* exy,
                    float gamma, float lambda, const T* u, const T* up, const T* gp, T* grad,
                    T* dbgForce, T* dbgWl, ReduceWs& ws, const Ctl* ctl, const BufTable<T>& bufs,
                    hipStream_t s) {
    const T invG = invGammaOf<T>(gamma);
    ws.blocks[3] = nodeBlocks(d.nCells, d.nOpt);
    launchKernel(k_node_grad<T>, ws.blocks[3], kBlock, 0, s, d, g, exy, invG, lambda, u, up, gp,
                 grad, dbgForce, dbgWl, ws.region(3), ctl, bufs);
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
void launchReduce(ReduceWs& ws, double denom, bool withNode, Ctl* ctl, DevRow* rows,
                  hipStream_t s) {
    launchKernel(k_reduce<T>, 1, kBlock, 0, s, ws, denom, withNode ? 1 : 0, ctl, rows);
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
void launchUpdate(const DevDesign<T>& d, const T* A, const T* G, const T* B, double alpha,
                  double m, T* vOut, T* uOut, const Ctl* ctl, const BufTable<T>& bufs,
                  hipStream_t s) {
    launchKernel(k_update<T>, nodeBlocks(d.nCells, d.nOpt), kBlock, 0, s, d, A, G, B,
                 static_cast<T>(alpha), static_cast<T>(m), vOut, uOut, ctl, bufs);
    SPHIP_CHECK(hipGetLastError());
}

#define SPHIP_INST(T)                                                                            \
    template int waResidentBlocks<T>();                                                          \
    template void launchScatter<T>(const DevDesign<T>&, const GridGeom<T>&, T,                   \
                                   unsigned long long*, unsigned long long*, bool, ReduceWs&,    \
                                   const Ctl*, hipStream_t);                                                 \
    template void launchTileSort<T>(const DevDesign<T>&, const GridGeom<T>&, const Ctl*,         \
                                    hipStream_t);                                                \
    template void launchFinalize<T>(const GridGeom<T>&, const TileGrid&, unsigned long long*,     \
                                    unsigned long long*, \
                                    const double*, const double*, double, double, double, T*,    \
                                    long long*, long long*, ReduceWs&, const Ctl*, hipStream_t); \
    template void launchWaNets<T>(const DevDesign<T>&, float, int, ReduceWs&, const Ctl*,        \
                                  hipStream_t);                                                  \
    template void launchNodeGrad<T>(const DevDesign<T>&, const GridGeom<T>&, const T*,           \
                                    float, float, const T*, const T*, const T*, T*, T*, T*,      \
                                    ReduceWs&, const Ctl*, const BufTable<T>&, hipStream_t);     \
    template void launchReduce<T>(ReduceWs&, double, bool, Ctl*, DevRow*, hipStream_t);         \
    template void launchUpdate<T>(const DevDesign<T>&, const T*, const T*, const T*, double,     \
                                  double, T*, T*, const Ctl*, const BufTable<T>&, hipStream_t);
SPHIP_INST(float)
SPHIP_INST(double)
template void launchSlotBatch<float>(int, const SlotArgs<float>*, const int4*, int, int,
                                     hipStream_t);

}  // namespace sphip
