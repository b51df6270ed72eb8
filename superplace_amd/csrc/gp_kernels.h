// Device-side data of one placement instance and launchers of the
// per-evaluation kernels (density scatter, rho/tau, WA per net, fused
// gather + WA node pass + gradient, reduction + decisions, Nesterov update).
#pragma once

#include "common.h"

namespace sphip {

// Spatial tiles of kTileBins x kTileBins bins.  Each tile owns a window of
// kTileWin x kTileWin bins (the tile and a halo of kTileHalo bins) in
// `stage`, which the scatter writes with plain stores and the finalize pass
// sums into the bins; the per-tile counters are zero between evaluations.
struct TileGrid {
    int tilesX = 0, tilesY = 0;   // tiles per grid side (the last ones may be partial)
    int nKeys = 0;                // Morton key range (a power of 4 >= tilesX * tilesY)
    int* count = nullptr;         // nodes per key (pass 1; cleared by the finalize pass)
    int* fill = nullptr;          // placement cursor per key (pass 2)
    int* start = nullptr;         // first record position of each key (last sort)
    int* size = nullptr;          // records of each key (last sort)
    int* extraStart = nullptr;    // nKeys + 1: first extra scatter block of each key
    unsigned long long* stage = nullptr;   // [tile][node | filler][kTileWin^2]
    int chunk = 512;              // records per scatter block (a tile's owner takes the first)
};

template <typename T>
struct DevDesign {
    int total = 0, numMov = 0, numNodes = 0;
    int nOpt = 0, nDens = 0, numNets = 0, numPins = 0;
    T coreLx, coreLy, coreHx, coreHy;
    // nodes
    T* xy = nullptr;         // current evaluation point of all nodes, (x, y) interleaved
    const T* w = nullptr;
    const T* h = nullptr;
    const uint8_t* flags = nullptr;
    const int* optIdx = nullptr;     // pos index k -> node i
    // When optIdx is k for k < optSplit and k + optShift beyond (the usual
    // opt set: movable nodes, then the fillers past the fixed nodes),
    // optAffine is set and the kernels compute i instead of loading it.
    int optAffine = 0, optSplit = 0, optShift = 0;
    int nCells = 0;        // optimised nodes that are not fillers (opt indices [0, nCells))
    int cellBlocks = 0;    // blocks of the cell part of the node-pass and update grids
    // The density nodes in spatial order (by tile of their centre bin, tiles
    // in Morton order) as records: node index with the filler / macro flags
    // in bits 30 / 29 (recId) and (w, h) interleaved (recWH).  The sort and
    // the scatter gather a record's position from xy through its node index
    // (a copy kept in record order cost the update more in scattered writes
    // than the gathers cost).  launchTileSort reads the previous order (A)
    // and writes the new one (B); the scatter copies B back to A.
    int* recIdA = nullptr;
    int* recIdB = nullptr;
    T* recWHA = nullptr;
    T* recWHB = nullptr;
    TileGrid tiles;
    // Nets are renumbered by degree (ascending), so each WA size class below
    // is a contiguous range of nets.  Classes below kWaClasses keep their
    // pins in lane order (waLaneBase; node -1: an empty slot), with their pin
    // offsets and node-pin entries (where their gradients go).
    const int* laneNode = nullptr;
    const T* laneOffX = nullptr;
    const T* laneOffY = nullptr;
    const int* laneNp = nullptr;
    // Tail nets (class kWaClasses): net waNetLo[kWaClasses] + k owns slots
    // [tailStart[k], tailStart[k + 1]) of the slot arrays.
    const int* tailStart = nullptr;
    const int* slotNode = nullptr;
    const T* slotOffX = nullptr;
    const T* slotOffY = nullptr;
    const int* slotNp = nullptr;
    // WA size classes: class c < kWaClasses holds the nets of degree in
    // (2^c, 2^(c+1)] (class 0: degree 2), one lane per net (2^(c+1) pin
    // slots, 64 nets per wavefront) below kWaLaneClasses, else 2^(c+1) lanes
    // per net; class kWaClasses the nets of degree > 2^kWaClasses, one
    // wavefront each.
    // Nets of class c: [waNetLo[c], waNetLo[c+1]); its wavefronts start at
    // waWaveBase[c]; waWaveBase[kWaClasses + 1] is the total.
    int waNetLo[8];
    int waWaveBase[8];
    // Wavefront tasks grouped by XCD: XCD x takes waTaskOrder[waXcdStart[x],
    // waXcdStart[x + 1]) (tasks whose nets start in its share of the nodes,
    // longest first), on its blocks b = x, x + kXcds, ... of the WA pass.
    const int* waTaskOrder = nullptr;
    int waXcdStart[kXcds + 1];
    int waBlocks = 1;     // blocks of every WA pass of the instance (fixes its partials' order)
    long waLaneBase[8];   // first lane-order entry of class c (c <= kWaClasses)
    // node -> pin entries; pinGrad holds the WA gradient (x, y) of each entry,
    // written by the net pass (zero for pins of nets it skips)
    const int* nodeStart = nullptr;
    T* pinGrad = nullptr;
};

constexpr int kWaClasses = 6;        // lanes per net 2, 4, ..., 64; then one wave per net
constexpr int kWaLaneClasses = 2;    // classes 0, 1 (degree 2, 3-4): one lane per net
// Nets of class c per wavefront task, and lane-order entries per task.
__host__ __device__ inline int waNetsPerTask(int c) { return c < kWaLaneClasses ? 64 : 64 >> (c + 1); }
__host__ __device__ inline int waEntriesPerTask(int c) { return c < kWaLaneClasses ? 64 * (2 << c) : 64; }
constexpr int kWaTailRegs = 8;       // pins per lane per batch of a tail net's extrema pass
// The node pass and the update run over the optimised nodes in two parts,
// the non-filler nodes (opt indices [0, m)) and the fillers ([m, n)), each
// part on a multiple of kXcds blocks; xcdBlock inside each part gives every
// XCD a contiguous share of the cells and one of the fillers.  The WA pass
// writes the pin entries of a share of the cells from the same XCD, so the
// node pass reads them from that XCD's L2, and every XCD has cells.
inline int nodePartBlocks(int count) { return roundUp(ceilDiv(count, kBlock), kXcds); }
inline int nodeBlocks(int m, int n) {
    return std::max(kXcds, nodePartBlocks(m) + nodePartBlocks(n - m));
}
constexpr int kWaMaxBlocks = 2048;   // the WA pass strides its wavefront tasks over at most these
// Blocks of the WA pass for `waves` wavefront tasks (kBlock threads each),
// at most `resident` (the blocks the GPU holds at once, so that no block
// waits for another to end before it starts), a multiple of kXcds (every
// XCD takes the tasks of its share of the nodes).
inline int waBlocksFor(int waves, int resident = kWaMaxBlocks) {
    const int b = roundUp(std::max(1, ceilDiv(waves, kBlock / 64)), kXcds);
    const int cap = std::max(kXcds, std::min(resident, kWaMaxBlocks) / kXcds * kXcds);
    return std::min(b, cap);
}
// Blocks of the WA pass's kernel (k_wa_pins) one GPU holds at once.
template <typename T>
int waResidentBlocks();
constexpr int kTileBins = 16;        // bins per tile side
constexpr int kTileHalo = 4;         // halo bins of a tile's scatter window
constexpr int kTileWin = kTileBins + 2 * kTileHalo;
constexpr int kRecFiller = 1 << 30, kRecMacro = 1 << 29, kRecNode = (1 << 29) - 1;
constexpr int kTileKeys = 4096;      // Morton keys of up to 64 x 64 tiles (1024^2 bins)
constexpr int kTileChunk = 512;      // default records per scatter block (TileGrid::chunk)
// The record chunks the scatter kernels are built for (TileGrid::chunk).
inline bool scatterChunkOk(int c) { return c == 256 || c == 512 || c == 1024 || c == 2048; }
constexpr double kSpillFrac = 0.05;  // re-sort once this fraction of nodes left their tile window
constexpr int kSortPeriod = 2;       // ... checked at every second evaluation of the device loop
constexpr int kSortBlocks = 512;     // grid of the sort passes (they stride over record chunks)

// FP64 reductions: every reducing kernel writes one partial per block (plain
// stores, no atomics, no fences); k_reduce sums each region in block order
// after the node pass.  Regions and their values per block:
//   0 finalize (tau numerator), 1 WA net pass (WL, HPWL), 2 scatter (nodes
//   with charge outside their tile window), 3 node pass (7 sums).
struct ReduceWs {
    double* partials = nullptr;   // 4 regions of maxBlocks * 8 doubles
    double* scalars = nullptr;    // EvalScalars on the device
    int maxBlocks = 0;
    int blocks[4] = {0, 0, 0, 0}; // blocks each region's kernel launches
    __host__ __device__ double* region(int r) const {
        return partials + static_cast<long>(r) * maxBlocks * 8;
    }
};

// ---------------------------------------------------------------------------
// Device-resident loop state.  Every iteration is the same fixed slot of
// kernels (scatter, finalize, transforms, WA, node pass, reduce, update); the
// data-dependent decisions of EPlace::run / NesterovOptimizer::step are taken
// by one thread of k_reduce and recorded here.
constexpr int kMaxSnaps = 4;      // tau-ladder snapshots (the ladder holds <= 3)
constexpr int kNumBufs = 8 + kMaxSnaps;
// roles in the state-buffer table (then kMaxSnaps snapshot buffers)
enum : int { B_U = 0, B_V, B_UP, B_VP, B_UN, B_VN, B_G, B_GP, B_ROLES };

// The state-buffer table, passed to the loop kernels by value: a role's
// buffer is then picked from kernel arguments by the index in Ctl, with no
// dependent load of the table itself.
template <typename T>
struct BufTable {
    T* p[kNumBufs];
};
template <typename T>
__host__ __device__ inline T* pickBuf(const BufTable<T>& t, int i) {
    T* r = t.p[0];
#pragma unroll
    for (int j = 1; j < kNumBufs; ++j) r = i == j ? t.p[j] : r;
    return r;
}

enum : int { kModeNone = 0, kModeStep = 1, kModeRedo = 2 };
enum : int { kRunning = 0, kConverged = 1, kStagnated = 2, kIterLimit = 3, kNonFinite = 4 };

struct DevRow {   // one metrics row (src/util/metrics.h:12-28)
    int iter, evaluations;
    double hpwl, overflow, lambda, gamma, step, gradWl, gradDen, evalLambda, evalGamma, dH, mu;
    unsigned long long ticks;   // device real-time clock when the row was written
};

// (The decisions keep every field but the tau-ladder arrays in registers:
// a field added here is added to SPHIP_CTL_SCALARS in gp_kernels.hip too.)
struct Ctl {
    // cur: table indices the update of this slot uses; nxt: the indices the
    // next evaluation uses (after this slot's step rotation).
    int cur[B_ROLES], nxt[B_ROLES];
    int mode;        // update of this slot: none / step / redo
    int done, status;
    int iter, tries, maxIter, maxBt, backtrack, useBB, stage, window;
    long long evals;
    int finalBuf, rollbackIter;
    int snapWrite, snapSrc;   // table index to copy the evaluated point into (-1 none)
    int nSnap, nFree;
    int snapBuf[kMaxSnaps], snapIter[kMaxSnaps], freeSnap[kMaxSnaps];
    float snapTau[kMaxSnaps];
    double snapHpwl[kMaxSnaps];
    double a, alpha, alpha0, m, momPrev;
    float lambda, gamma;      // schedule values used by the next evaluation
    float targetTau, stepX, minGain, bestTau, refTau;
    int sinceImprove, havePrev;
    int sortNext;             // the next evaluation re-sorts the density records
    double spillLimit;        // ... when more nodes than this left their tile window
    double prevHpwl, effRef;
};

// Every launcher takes an optional device loop state `ctl` (null in the
// host-driven mode).  With ctl, kernels exit at once when ctl->done is set,
// gamma/lambda come from ctl, and the node pass / update resolve their state
// vectors through the table `bufs` (kNumBufs device pointers).
// Density scatter in tile order (after launchTileSort): one owner block per
// tile accumulates the fixed-point charges of the tile's first kTileChunk
// nodes in its window and stores the window to tiles.stage; extra blocks take
// the further chunks of crowded tiles and add their windows to nodeQ / fillQ
// with atomics, as do charges that fall outside a window (large nodes).
// `sorted` (with ctl: and ctl->sortNext): launchTileSort ran for this
// evaluation, so the scatter reads the new order (B) and copies it to A;
// otherwise it reads A, the order of the last sort, gathering the current
// positions (charges of nodes that left their tile's window
// since then go through the atomic grids).  The count of such nodes per
// block goes to region 2 of ws; k_reduce re-sorts once it passes
// ctl->spillLimit.
template <typename T>
void launchScatter(const DevDesign<T>& d, const GridGeom<T>& g, T qscale,
                   unsigned long long* nodeQ, unsigned long long* fillQ, bool sorted,
                   ReduceWs& ws, const Ctl* ctl, hipStream_t s);

// Re-sorts the density nodes by tile (counting sort; the order inside a tile
// is arbitrary, which the integer density sums do not see).  With ctl, the
// passes return at once unless ctl->sortNext is set.
template <typename T>
void launchTileSort(const DevDesign<T>& d, const GridGeom<T>& g, const Ctl* ctl, hipStream_t s);


// Sums the staged windows and the atomic grids into node / filler density,
// then rho and the tau numerator; clears the atomic grids and tile counters.
template <typename T>
void launchFinalize(const GridGeom<T>& g, const TileGrid& tg, unsigned long long* nodeQ,
                    unsigned long long* fillQ, const double* terminal, const double* base,
                    double invQ, double binArea, double rhoShift, T* rho, long long* dbgNode,
                    long long* dbgFill, ReduceWs& ws, const Ctl* ctl, hipStream_t s);

template <typename T>
void launchWaNets(const DevDesign<T>& d, float gamma, int ignoreDeg, ReduceWs& ws,
                  const Ctl* ctl, hipStream_t s);

template <typename T>
void launchNodeGrad(const DevDesign<T>& d, const GridGeom<T>& g, const T* exy,
                    float gamma, float lambda, const T* u, const T* up, const T* gp, T* grad,
                    T* dbgForce, T* dbgWl, ReduceWs& ws, const Ctl* ctl, const BufTable<T>& bufs,
                    hipStream_t s);

// Sums the regions written by this evaluation into EvalScalars (tau is the
// finalize sum over `denom`); withNode = false for a density + HPWL pass only.
// With ctl, one thread then takes the step decisions and writes `rows`.
template <typename T>
void launchReduce(ReduceWs& ws, double denom, bool withNode, Ctl* ctl, DevRow* rows,
                  hipStream_t s);

// v' = A - alpha G; u' = v' + m (v' - B); clamp both to the core; write
// vOut/uOut (either may be null) and scatter u' (or v' if uOut is null) into
// the node coordinates.  G null means 0, B null means v'.  With ctl, the
// operands, alpha and m come from ctl->mode (step or redo) and the table.
template <typename T>
void launchUpdate(const DevDesign<T>& d, const T* A, const T* G, const T* B, double alpha,
                  double m, T* vOut, T* uOut, const Ctl* ctl, const BufTable<T>& bufs,
                  hipStream_t s);

}  // namespace sphip
