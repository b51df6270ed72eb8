// Device-side data of one placement instance and launchers of the
// per-evaluation kernels (density scatter, rho/tau, WA per net, fused
// gather + WA node pass + gradient, Nesterov update).
#pragma once

#include "common.h"

namespace sphip {

// Per-net WA state written by pass 1 and read by the node pass
// (wa_wirelength.cpp:130-167): extrema, b+-, centred c'+- for x and y.
template <typename T>
struct alignas(16) NetStat {
    T mxX, mnX, mxY, mnY, bPX, bNX, cPX, cNX, bPY, bNY, cPY, cNY;
};

template <typename T>
struct DevDesign {
    int total = 0, numMov = 0, numNodes = 0;
    int nOpt = 0, nDens = 0, numNets = 0, numPins = 0;
    int nSmallNets = 0, nLargeNets = 0;
    T coreLx, coreLy, coreHx, coreHy;
    // nodes
    T* x = nullptr;          // current evaluation point, all nodes
    T* y = nullptr;
    const T* w = nullptr;
    const T* h = nullptr;
    const uint8_t* flags = nullptr;
    const int* optIdx = nullptr;     // pos index k -> node i
    const int* densIdx = nullptr;    // nodes with a density contribution
    const double* precArea = nullptr;  // area of opt node k (double)
    const float* precDeg = nullptr;    // pin count of opt node k
    // net -> pin slots (net order)
    const int* netStart = nullptr;
    const int* slotNode = nullptr;
    const T* slotOffX = nullptr;
    const T* slotOffY = nullptr;
    const int* smallNets = nullptr;   // nets of degree <= kLargeNet (thread per net)
    const int* largeNets = nullptr;   // nets of degree > kLargeNet (wave per net)
    // node -> pin entries
    const int* nodeStart = nullptr;
    const int* npNet = nullptr;       // net of each node-pin entry (-1 = none)
    const T* npOffX = nullptr;
    const T* npOffY = nullptr;
    NetStat<T>* stats = nullptr;
};

constexpr int kLargeNet = 32;

// Reduction workspace: per-block partials, counters and the final scalars.
struct ReduceWs {
    double* partials = nullptr;   // [slots]
    unsigned* counters = nullptr; // [8]
    double* scalars = nullptr;    // EvalScalars on the device
    int maxBlocks = 0;
};

template <typename T>
void launchScatter(const DevDesign<T>& d, const GridGeom<T>& g, T qscale,
                   unsigned long long* nodeQ, unsigned long long* fillQ, hipStream_t s);

template <typename T>
void launchFinalize(const GridGeom<T>& g, unsigned long long* nodeQ, unsigned long long* fillQ,
                    const double* terminal, const double* base, double invQ, double binArea,
                    double movableAreaScaled, double rhoShift, T* rho, long long* dbgNode,
                    long long* dbgFill, const ReduceWs& ws, hipStream_t s);

template <typename T>
void launchWaNets(const DevDesign<T>& d, float gamma, int ignoreDeg, const ReduceWs& ws,
                  hipStream_t s);

template <typename T>
void launchNodeGrad(const DevDesign<T>& d, const GridGeom<T>& g, const T* ex, const T* ey,
                    float gamma, float lambda, const T* u, const T* up, const T* gp, T* grad,
                    T* dbgForce, T* dbgWl, const ReduceWs& ws, hipStream_t s);

// v' = A - alpha G; u' = v' + m (v' - B); clamp both to the core; write
// vOut/uOut (either may be null) and scatter u' (or v' if uOut is null) into
// the node coordinates.  G null means 0, B null means v'.
template <typename T>
void launchUpdate(const DevDesign<T>& d, const T* A, const T* G, const T* B, double alpha,
                  double m, T* vOut, T* uOut, hipStream_t s);

}  // namespace sphip
