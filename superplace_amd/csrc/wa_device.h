// Device code of the WA wirelength pass and the FP64 block reductions,
// shared by the per-evaluation kernels and by the column pass of the
// Poisson solve, which can carry the WA pass's blocks in its own launch
// (the two are independent: WA reads only positions).
#pragma once

#include "gp_kernels.h"

namespace sphip {

// ------------------------------------------------------------------ reductions
template <int NV>
__device__ inline void blockSum(double (&v)[NV]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += __shfl_down(v[j], off, 64);
    __shared__ double sh[kBlock / 64][NV];
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) sh[wave][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double s = 0;
            for (int w = 0; w < kBlock / 64; ++w) s += sh[w][j];
            v[j] = s;
        }
    __syncthreads();
}

// One partial per block: partials[block * NV + j] (block defaults to blockIdx.x).
template <int NV>
__device__ inline void blockPartial(double (&v)[NV], double* partials, int block = -1) {
    blockSum<NV>(v);
    if (block < 0) block = blockIdx.x;
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) partials[static_cast<long>(block) * NV + j] = v[j];
}

// ------------------------------------------------------------------ (c) WA pass 1 + HPWL
template <typename T>
__device__ inline T texp(T v);
template <>
__device__ inline float texp<float>(float v) { return expf(v); }
template <>
__device__ inline double texp<double>(double v) { return exp(v); }

// wa_wirelength.cpp:95-97: gamma floored at 1e-4, inverted in the working precision.
template <typename T>
__device__ __host__ inline T invGammaOf(float gamma) {
    const T g = static_cast<T>(gamma) > T(1e-4f) ? static_cast<T>(gamma) : T(1e-4f);
    return T(1) / g;
}

template <typename T>
__device__ inline T tmax() { return sizeof(T) == 4 ? T(3.402823466e+38f) : T(1.7976931348623157e308); }

template <typename T>
__device__ inline T waveMax(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
template <typename T>
__device__ inline T waveMin(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
template <typename T>
__device__ inline T waveSum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// Reductions over aligned groups of W lanes (xor offsets stay in the group).
template <int W, typename T>
__device__ inline T groupMax(T v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
template <int W, typename T>
__device__ inline T groupMin(T v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
template <int W, typename T>
__device__ inline T groupSum(T v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sums of one net over its pins, after the extrema (wa_wirelength.cpp:130-167):
// b+- = sum a+-, c'+- = sum (p - extremum) a+-, per axis.
template <typename T>
struct WaSums {
    T bpx, bnx, cpx, cnx, bpy, bny, cpy, cny;
};

// WA gradient of one pin from its exponentials and the net's sums: the
// derivative of the max term minus that of the min term
// (wa_wirelength.cpp:198-218), both axes.
template <typename T>
__device__ inline void pinGradOf(const WaSums<T>& s, T dxp, T dxn, T ax, T anx, T dyp, T dyn,
                                 T ay, T any, T invG, T& gx, T& gy) {
    const T tMaxX = ax * (s.bpx + (dxp * s.bpx - s.cpx) * invG) / (s.bpx * s.bpx);
    const T tMinX = anx * (s.bnx + (s.cnx - dxn * s.bnx) * invG) / (s.bnx * s.bnx);
    const T tMaxY = ay * (s.bpy + (dyp * s.bpy - s.cpy) * invG) / (s.bpy * s.bpy);
    const T tMinY = any * (s.bny + (s.cny - dyn * s.bny) * invG) / (s.bny * s.bny);
    gx = tMaxX - tMinX;
    gy = tMaxY - tMinY;
}

template <typename T>
__device__ inline double waOf(T mxX, T mnX, T mxY, T mnY, const WaSums<T>& s) {
    return static_cast<double>((mxX - mnX) + s.cpx / s.bpx - s.cnx / s.bnx) +
           static_cast<double>((mxY - mnY) + s.cpy / s.bpy - s.cny / s.bny);
}

// Size class of nets of at most P pins with one lane per net: the lane
// holds its net's pins in registers, so extrema, sums and gradients need no
// cross-lane step, and the sums run over the pins in net order as the
// reference's loops do (wa_wirelength.cpp:119-225).  Pin j of the lane's net
// is lane-order entry (task P + j) 64 + lane: every load is coalesced and
// needs no per-net index first.  Empty slots (node -1) carry no pin.
template <int P, typename T>
__device__ inline void waLanePass(const DevDesign<T>& d, int c, int gw, int lane, T invG,
                                  int ignoreDeg, double (&v)[2]) {
    const long e0 = d.waLaneBase[c] + static_cast<long>(gw - d.waWaveBase[c]) * P * 64 + lane;
    int nd[P], q[P];
    T px[P], py[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const long e = e0 + 64L * j;
        nd[j] = d.laneNode[e];
        px[j] = d.laneOffX[e];
        py[j] = d.laneOffY[e];
        q[j] = d.laneNp[e];
    }
    int deg = 0;
#pragma unroll
    for (int j = 0; j < P; ++j)
        if (nd[j] >= 0) {
            const auto p = load2(d.xy, nd[j]);
            px[j] = p.x + px[j];
            py[j] = p.y + py[j];
            ++deg;
        }
    if (deg < 2) return;
    T mxX = -tmax<T>(), mnX = tmax<T>(), mxY = -tmax<T>(), mnY = tmax<T>();
#pragma unroll
    for (int j = 0; j < P; ++j)
        if (nd[j] >= 0) {
            mxX = fmax(mxX, px[j]); mnX = fmin(mnX, px[j]);
            mxY = fmax(mxY, py[j]); mnY = fmin(mnY, py[j]);
        }
    v[1] += static_cast<double>(mxX - mnX) + static_cast<double>(mxY - mnY);
    if (deg > ignoreDeg) return;
    WaSums<T> s{0, 0, 0, 0, 0, 0, 0, 0};
    T ax[P], anx[P], ay[P], any[P];
#pragma unroll
    for (int j = 0; j < P; ++j)
        if (nd[j] >= 0) {
            const T dxp = px[j] - mxX, dxn = px[j] - mnX, dyp = py[j] - mxY, dyn = py[j] - mnY;
            ax[j] = texp<T>(dxp * invG); anx[j] = texp<T>(-dxn * invG);
            ay[j] = texp<T>(dyp * invG); any[j] = texp<T>(-dyn * invG);
            s.bpx += ax[j]; s.cpx += dxp * ax[j]; s.bnx += anx[j]; s.cnx += dxn * anx[j];
            s.bpy += ay[j]; s.cpy += dyp * ay[j]; s.bny += any[j]; s.cny += dyn * any[j];
        }
    v[0] += waOf<T>(mxX, mnX, mxY, mnY, s);
#pragma unroll
    for (int j = 0; j < P; ++j)
        if (nd[j] >= 0) {
            const T dxp = px[j] - mxX, dxn = px[j] - mnX, dyp = py[j] - mxY, dyn = py[j] - mnY;
            T gx, gy;
            pinGradOf<T>(s, dxp, dxn, ax[j], anx[j], dyp, dyn, ay[j], any[j], invG, gx, gy);
            store2(d.pinGrad, q[j], gx, gy);
        }
}

// Size class with W = 2^LW lanes per net: one pin per lane, so a pin is
// loaded once and its exponentials are computed once, for both the net's
// sums and the pin's gradient.  The class's pins sit in lane order (entry
// waLaneBase[c] + 64 task + lane; node -1 marks an empty lane), so the
// lane's node, offsets and node-pin entry are one coalesced load with no
// per-net index before it.  Empty lanes carry neutral values.
template <int LW, typename T>
__device__ inline void waGroupPass(const DevDesign<T>& d, int c, int gw, int lane, T invG,
                                   int ignoreDeg, double (&v)[2]) {
    constexpr int W = 1 << LW;
    const long e = d.waLaneBase[c] + static_cast<long>(gw - d.waWaveBase[c]) * 64 + lane;
    const int j = lane & (W - 1);
    const int node = d.laneNode[e];
    const T ox = d.laneOffX[e], oy = d.laneOffY[e];
    const int q = d.laneNp[e];
    const bool pin = node >= 0;
    T px = 0, py = 0;
    if (pin) {
        const auto p = load2(d.xy, node);
        px = p.x + ox;
        py = p.y + oy;
    }
    const int deg = groupSum<W>(pin ? 1 : 0);
    const T mxX = groupMax<W>(pin ? px : -tmax<T>()), mnX = groupMin<W>(pin ? px : tmax<T>());
    const T mxY = groupMax<W>(pin ? py : -tmax<T>()), mnY = groupMin<W>(pin ? py : tmax<T>());
    const T dxp = px - mxX, dxn = px - mnX, dyp = py - mxY, dyn = py - mnY;
    const T ax = pin ? texp<T>(dxp * invG) : T(0), anx = pin ? texp<T>(-dxn * invG) : T(0);
    const T ay = pin ? texp<T>(dyp * invG) : T(0), any = pin ? texp<T>(-dyn * invG) : T(0);
    WaSums<T> s;
    s.bpx = groupSum<W>(ax); s.cpx = groupSum<W>(dxp * ax);
    s.bnx = groupSum<W>(anx); s.cnx = groupSum<W>(dxn * anx);
    s.bpy = groupSum<W>(ay); s.cpy = groupSum<W>(dyp * ay);
    s.bny = groupSum<W>(any); s.cny = groupSum<W>(dyn * any);
    const bool wa = deg >= 2 && deg <= ignoreDeg;
    if (j == 0 && deg >= 2) {
        v[1] += static_cast<double>(mxX - mnX) + static_cast<double>(mxY - mnY);
        if (wa) v[0] += waOf<T>(mxX, mnX, mxY, mnY, s);
    }
    if (wa && pin) {
        T gx, gy;
        pinGradOf<T>(s, dxp, dxn, ax, anx, dyp, dyn, ay, any, invG, gx, gy);
        store2(d.pinGrad, q, gx, gy);
    }
}

// A tail net of more than kWaTailRegs * 64 pins: lane-strided loops over
// its slots (extrema, sums, gradients), the exponentials recomputed in the
// gradient loop.
template <typename T>
__device__ inline void waWaveLoop(const DevDesign<T>& d, int b, int e, int lane, T invG,
                                  int ignoreDeg, double (&v)[2]) {
    T mxX = -tmax<T>(), mnX = tmax<T>(), mxY = -tmax<T>(), mnY = tmax<T>();
    for (int t = b + lane; t < e; t += 64) {
        const auto p = load2(d.xy, d.slotNode[t]);
        const T px = p.x + d.slotOffX[t], py = p.y + d.slotOffY[t];
        mxX = fmax(mxX, px); mnX = fmin(mnX, px);
        mxY = fmax(mxY, py); mnY = fmin(mnY, py);
    }
    mxX = waveMax(mxX); mnX = waveMin(mnX); mxY = waveMax(mxY); mnY = waveMin(mnY);
    if (lane == 0) v[1] += static_cast<double>(mxX - mnX) + static_cast<double>(mxY - mnY);
    if (e - b > ignoreDeg) return;
    WaSums<T> s{0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = b + lane; t < e; t += 64) {
        const auto p = load2(d.xy, d.slotNode[t]);
        const T px = p.x + d.slotOffX[t], py = p.y + d.slotOffY[t];
        const T dxp = px - mxX, dxn = px - mnX, dyp = py - mxY, dyn = py - mnY;
        const T ax = texp<T>(dxp * invG), anx = texp<T>(-dxn * invG);
        const T ay = texp<T>(dyp * invG), any = texp<T>(-dyn * invG);
        s.bpx += ax; s.cpx += dxp * ax; s.bnx += anx; s.cnx += dxn * anx;
        s.bpy += ay; s.cpy += dyp * ay; s.bny += any; s.cny += dyn * any;
    }
    s.bpx = waveSum(s.bpx); s.cpx = waveSum(s.cpx); s.bnx = waveSum(s.bnx); s.cnx = waveSum(s.cnx);
    s.bpy = waveSum(s.bpy); s.cpy = waveSum(s.cpy); s.bny = waveSum(s.bny); s.cny = waveSum(s.cny);
    if (lane == 0) v[0] += waOf<T>(mxX, mnX, mxY, mnY, s);
    for (int t = b + lane; t < e; t += 64) {
        const auto p = load2(d.xy, d.slotNode[t]);
        const T px = p.x + d.slotOffX[t], py = p.y + d.slotOffY[t];
        const T dxp = px - mxX, dxn = px - mnX, dyp = py - mxY, dyn = py - mnY;
        const T ax = texp<T>(dxp * invG), anx = texp<T>(-dxn * invG);
        const T ay = texp<T>(dyp * invG), any = texp<T>(-dyn * invG);
        T gx, gy;
        pinGradOf<T>(s, dxp, dxn, ax, anx, dyp, dyn, ay, any, invG, gx, gy);
        const int q = d.slotNp[t];
        store2(d.pinGrad, q, gx, gy);
    }
}

// Nets above the largest class (the power-law tail): one wavefront per net,
// lane l taking the net's pins l, l + 64, ...  The slot loads of a batch of
// pins are issued together, then its coordinate gathers together, so a batch
// costs two memory round trips:
//  - a net above ignoreDeg needs only its extrema (HPWL): batches of
//    kWaTailRegs pins per lane, folded into the extrema as they arrive;
//  - a net of at most 2 x 64 pins (within ignoreDeg) keeps its pins in
//    registers for the extrema, the sums and the gradients;
//  - larger nets within ignoreDeg run waWaveLoop.
// Extrema do not depend on order, and the sums run over a lane's pins in the
// order of waWaveLoop's loops, so every path gives waWaveLoop's bits.
template <typename T>
__device__ inline void waWavePass(const DevDesign<T>& d, int gw, int lane, T invG, int ignoreDeg,
                                  double (&v)[2]) {
    const int k = gw - d.waWaveBase[kWaClasses];
    if (d.waNetLo[kWaClasses] + k >= d.waNetLo[kWaClasses + 1]) return;
    const int b = d.tailStart[k], e = d.tailStart[k + 1];
    if (e - b > ignoreDeg) {
        T mxX = -tmax<T>(), mnX = tmax<T>(), mxY = -tmax<T>(), mnY = tmax<T>();
        constexpr int R = kWaTailRegs;
        for (int t0 = b + lane; t0 < e; t0 += 64 * R) {
            int nd[R];
            T ox[R], oy[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int t = t0 + 64 * r;
                const bool in = t < e;
                nd[r] = in ? d.slotNode[t] : -1;
                ox[r] = in ? d.slotOffX[t] : T(0);
                oy[r] = in ? d.slotOffY[t] : T(0);
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (nd[r] >= 0) {
                    const auto p = load2(d.xy, nd[r]);
                    const T px = p.x + ox[r], py = p.y + oy[r];
                    mxX = fmax(mxX, px); mnX = fmin(mnX, px);
                    mxY = fmax(mxY, py); mnY = fmin(mnY, py);
                }
        }
        mxX = waveMax(mxX); mnX = waveMin(mnX); mxY = waveMax(mxY); mnY = waveMin(mnY);
        if (lane == 0) v[1] += static_cast<double>(mxX - mnX) + static_cast<double>(mxY - mnY);
        return;
    }
    if (e - b > 128) {
        waWaveLoop<T>(d, b, e, lane, invG, ignoreDeg, v);
        return;
    }
    constexpr int R = 2;
    int nd[R], q[R];
    T px[R], py[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int t = b + lane + 64 * r;
        const bool in = t < e;
        nd[r] = in ? d.slotNode[t] : -1;
        px[r] = in ? d.slotOffX[t] : T(0);
        py[r] = in ? d.slotOffY[t] : T(0);
        q[r] = in ? d.slotNp[t] : 0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (nd[r] >= 0) {
            const auto p = load2(d.xy, nd[r]);
            px[r] = p.x + px[r];
            py[r] = p.y + py[r];
        }
    T mxX = -tmax<T>(), mnX = tmax<T>(), mxY = -tmax<T>(), mnY = tmax<T>();
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (nd[r] >= 0) {
            mxX = fmax(mxX, px[r]); mnX = fmin(mnX, px[r]);
            mxY = fmax(mxY, py[r]); mnY = fmin(mnY, py[r]);
        }
    mxX = waveMax(mxX); mnX = waveMin(mnX); mxY = waveMax(mxY); mnY = waveMin(mnY);
    if (lane == 0) v[1] += static_cast<double>(mxX - mnX) + static_cast<double>(mxY - mnY);
    WaSums<T> s{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (nd[r] >= 0) {
            const T dxp = px[r] - mxX, dxn = px[r] - mnX, dyp = py[r] - mxY, dyn = py[r] - mnY;
            const T ax = texp<T>(dxp * invG), anx = texp<T>(-dxn * invG);
            const T ay = texp<T>(dyp * invG), any = texp<T>(-dyn * invG);
            s.bpx += ax; s.cpx += dxp * ax; s.bnx += anx; s.cnx += dxn * anx;
            s.bpy += ay; s.cpy += dyp * ay; s.bny += any; s.cny += dyn * any;
        }
    s.bpx = waveSum(s.bpx); s.cpx = waveSum(s.cpx); s.bnx = waveSum(s.bnx); s.cnx = waveSum(s.cnx);
    s.bpy = waveSum(s.bpy); s.cpy = waveSum(s.cpy); s.bny = waveSum(s.bny); s.cny = waveSum(s.cny);
    if (lane == 0) v[0] += waOf<T>(mxX, mnX, mxY, mnY, s);
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (nd[r] >= 0) {
            const T dxp = px[r] - mxX, dxn = px[r] - mnX, dyp = py[r] - mxY, dyn = py[r] - mnY;
            const T ax = texp<T>(dxp * invG), anx = texp<T>(-dxn * invG);
            const T ay = texp<T>(dyp * invG), any = texp<T>(-dyn * invG);
            T gx, gy;
            pinGradOf<T>(s, dxp, dxn, ax, anx, dyp, dyn, ay, any, invG, gx, gy);
            store2(d.pinGrad, q[r], gx, gy);
        }
}


// The WA pass's share of one block: wavefront tasks block * waves ..
// strided over nblocks blocks of kBlock threads, the block's FP64 partials
// [WA, HPWL] written to partials[2 block].  Every launch with the same
// nblocks sums in the same order.
template <typename T>
struct WaTask {
    DevDesign<T> d;
    T invG;
    int ignoreDeg;
    double* partials;
    const Ctl* ctl;
};

template <typename T>
__device__ inline void waBlockTask(const WaTask<T>& w, int block, int nblocks) {
    T invG = w.invG;
    if (w.ctl) invG = invGammaOf<T>(w.ctl->gamma);
    const DevDesign<T>& d = w.d;
    const int ignoreDeg = w.ignoreDeg;
    const int lane = threadIdx.x & 63;
    double v[2] = {0.0, 0.0};
    // wavefront tasks are strided over the blocks (a bounded number of block
    // partials); each task belongs to one size class
    // Block b serves XCD b % kXcds (the XCD that runs it in a launch of its
    // own or behind a multiple of kXcds other blocks): the tasks of the nets
    // that start in that XCD's share of the nodes, longest first, so the
    // pin entries of a share are written from one L2 and the node pass,
    // which gives each XCD the same share, reads them there.
    const int x = block % kXcds, nbx = nblocks / kXcds;
    const int hi = d.waXcdStart[x + 1];
    for (int q = d.waXcdStart[x] + (block / kXcds) * (kBlock / 64) + (threadIdx.x >> 6); q < hi;
         q += nbx * (kBlock / 64)) {
        const int gw = d.waTaskOrder[q];
        int c = 0;
        while (c <= kWaClasses && gw >= d.waWaveBase[c + 1]) ++c;
        static_assert(kWaLaneClasses == 2, "classes 0 and 1 take one lane per net");
        switch (c) {
            case 0: waLanePass<2, T>(d, 0, gw, lane, invG, ignoreDeg, v); break;
            case 1: waLanePass<4, T>(d, 1, gw, lane, invG, ignoreDeg, v); break;
            case 2: waGroupPass<3, T>(d, 2, gw, lane, invG, ignoreDeg, v); break;
            case 3: waGroupPass<4, T>(d, 3, gw, lane, invG, ignoreDeg, v); break;
            case 4: waGroupPass<5, T>(d, 4, gw, lane, invG, ignoreDeg, v); break;
            case 5: waGroupPass<6, T>(d, 5, gw, lane, invG, ignoreDeg, v); break;
            case kWaClasses: waWavePass<T>(d, gw, lane, invG, ignoreDeg, v); break;
            default: break;
        }
    }
    blockPartial<2>(v, w.partials, block);
}

}  // namespace sphip
