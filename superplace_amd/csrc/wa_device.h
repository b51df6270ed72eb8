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

// Size class with W = 2^LW lanes per net: one pin per lane, so a pin is
// loaded once and its exponentials are computed once, for both the net's
// sums and the pin's gradient.  Lanes past the degree carry neutral values.
template <int LW, typename T>
__device__ inline void waGroupPass(const DevDesign<T>& d, int c, int gw, int lane, T invG,
                                   int ignoreDeg, double (&v)[2]) {
    constexpr int W = 1 << LW;
    const int net = d.waNetLo[c] + ((gw - d.waWaveBase[c]) << (6 - LW)) + (lane >> LW);
    const int j = lane & (W - 1);
    int t = 0, deg = 0;
    if (net < d.waNetLo[c + 1]) {
        t = d.netStart[net];
        deg = d.netStart[net + 1] - t;
        t += j;
    }
    const bool pin = j < deg;
    T px = 0, py = 0;
    if (pin) {
        const auto p = load2(d.xy, d.slotNode[t]);
        px = p.x + d.slotOffX[t];
        py = p.y + d.slotOffY[t];
    }
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
        const int q = d.slotNp[t];
        d.pinGrad[2L * q] = gx;
        d.pinGrad[2L * q + 1] = gy;
    }
}

// Nets above the largest class (the power-law tail): one wavefront per net,
// lane-strided pins, shuffle trees; the gradient loop recomputes the pins'
// exponentials.
template <typename T>
__device__ inline void waWavePass(const DevDesign<T>& d, int gw, int lane, T invG, int ignoreDeg,
                                  double (&v)[2]) {
    const int net = d.waNetLo[kWaClasses] + (gw - d.waWaveBase[kWaClasses]);
    if (net >= d.waNetLo[kWaClasses + 1]) return;
    const int b = d.netStart[net], e = d.netStart[net + 1];
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
        d.pinGrad[2L * q] = gx;
        d.pinGrad[2L * q + 1] = gy;
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
    const int tasks = d.waWaveBase[kWaClasses + 1];
    for (int gw = block * (kBlock / 64) + (threadIdx.x >> 6); gw < tasks;
         gw += nblocks * (kBlock / 64)) {
        int c = 0;
        while (c <= kWaClasses && gw >= d.waWaveBase[c + 1]) ++c;
        switch (c) {
            case 0: waGroupPass<1, T>(d, 0, gw, lane, invG, ignoreDeg, v); break;
            case 1: waGroupPass<2, T>(d, 1, gw, lane, invG, ignoreDeg, v); break;
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
