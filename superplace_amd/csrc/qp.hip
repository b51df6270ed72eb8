// Quadratic initial placement on the GPU: the Kraftwerk2 Bound2Bound model
// solved by BiCGSTAB, QuadraticPlacer::place (src/init/quadratic_placer.cpp:149-228).
//
// Per outer iteration, everything on the device:
//  1. k_qp_bounds   bound pins of every net (computeNetBoundPins,
//                   src/db/hpwl.cpp:70-102: first pin holding the min / max);
//  2. k_qp_count,   the B2B system of both axes as one CSR of 2 n rows
//     scan,         (row a n + i: movable node i on axis a), built row by row
//     k_qp_fill     in a fixed order (pins of the node, partners in net
//                   order), diagonal and right-hand side summed in FP64; the
//                   pairs are those of forEachB2BPair (src/init/b2b_model.h:56-82);
//  3. BiCGSTAB      right-preconditioned (Jacobi), both axes in the same
//                   launches, FP64 vectors; dot products are block partials
//                   summed in block order by the last block to finish, so
//                   every run gives the same bits; per-axis scalars and a
//                   done flag live on the device, so the qp_solver_max_iter
//                   iterations are enqueued without host round trips;
//  4. k_qp_clamp    positions back into the core (quadratic_placer.cpp:120-128);
//  5. k_qp_hpwl     HPWL for the outer stop rule.
// The host reads two scalars per outer iteration (HPWL, solver errors).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

#include "abi.h"

// No fused multiply-add anywhere in this file: the assembled systems equal a
// plain sequential evaluation bit for bit, and the solver's element-wise
// updates round like the oracle's.
#pragma clang fp contract(off)

namespace sphip {
namespace {

constexpr int kQpBlock = 256;
constexpr int kQpMaxBlocks = 1024;   // blocks per axis for the vector kernels

struct QpDesign {
    int nMov, nNodes, nPins, nNets;
    const float* w;
    const float* h;
    const int32_t* pin2node;
    const int32_t* pin2net;
    const float* off[2];      // pin offsets per axis
    const int32_t* net2pin;
    const int32_t* flatNet;
    const int32_t* node2pin;
    const int32_t* flatNode;
    const float* netW;        // may be null (all 1)
    float* pos[2];            // node positions per axis (all nodes)
    int32_t* bnd;             // 4 x nNets: min x, max x, min y, max y
    float core[4];            // lx, ly, hx, hy
};

// Per-axis BiCGSTAB state.
struct QpScal {
    double rho, rhoPrev, alpha, w, bnorm2, rr, tol2;
    int iter, done;
};

struct QpVec {
    double *x, *r, *r0, *p, *v, *y, *s, *z, *t, *b, *diag;
    const int64_t* rowStart;
    const int32_t* col;
    const double* val;
    int n;                    // rows per axis
    QpScal* sc;               // [2]
    double* part;             // [2][kQpMaxBlocks][4] block partials
    unsigned* count;          // [2] blocks finished, per axis
    int maxIter;
};

// ---- bound pins -----------------------------------------------------------
__global__ void k_qp_bounds(QpDesign d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.nNets) return;
    const int b = d.net2pin[k], e = d.net2pin[k + 1];
    int32_t* out = d.bnd;
    if (b == e) {
        out[k] = out[d.nNets + k] = out[2 * d.nNets + k] = out[3 * d.nNets + k] = -1;
        return;
    }
    const int p0 = d.flatNet[b];
    int a0 = p0, a1 = p0, a2 = p0, a3 = p0;
    const int n0 = d.pin2node[p0];
    float lox = d.pos[0][n0] + d.off[0][p0], hix = lox;
    float loy = d.pos[1][n0] + d.off[1][p0], hiy = loy;
    for (int t = b + 1; t < e; ++t) {
        const int p = d.flatNet[t];
        const int n = d.pin2node[p];
        const float px = d.pos[0][n] + d.off[0][p], py = d.pos[1][n] + d.off[1][p];
        if (px < lox) { lox = px; a0 = p; }
        if (px > hix) { hix = px; a1 = p; }
        if (py < loy) { loy = py; a2 = p; }
        if (py > hiy) { hiy = py; a3 = p; }
    }
    out[k] = a0;
    out[d.nNets + k] = a1;
    out[2 * d.nNets + k] = a2;
    out[3 * d.nNets + k] = a3;
}

// Visits the B2B partners of pin p on axis a in net order: f(q, w).
template <typename F>
__device__ inline void forPartners(const QpDesign& d, int a, int p, float minD, int ignoreDeg,
                                   F&& f) {
    const int k = d.pin2net[p];
    const int b = d.net2pin[k], e = d.net2pin[k + 1];
    const int deg = e - b;
    if (deg < 2 || deg > ignoreDeg) return;
    const int bmin = d.bnd[2 * a * d.nNets + k], bmax = d.bnd[(2 * a + 1) * d.nNets + k];
    const float coef = (d.netW ? d.netW[k] : 1.f) / static_cast<float>(deg - 1);
    const int np = d.pin2node[p];
    const float* pos = d.pos[a];
    const float* off = d.off[a];
    const float pp = pos[np] + off[p];
    const bool pBound = p == bmin || p == bmax;
    for (int t = b; t < e; ++t) {
        const int q = d.flatNet[t];
        if (q == p) continue;
        const int nq = d.pin2node[q];
        if (nq == np) continue;
        if (!pBound && q != bmin && q != bmax) continue;
        const float dist = fabsf(pp - (pos[nq] + off[q]));
        f(q, nq, coef / fmaxf(dist, minD));
    }
}

// Row r = a n + i: number of movable partners (off-diagonal entries).
__global__ void k_qp_count(QpDesign d, float minD, int ignoreDeg, int64_t* cnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 2 * d.nMov) return;
    const int a = r >= d.nMov, i = r - a * d.nMov;
    int64_t c = 0;
    for (int s = d.node2pin[i]; s < d.node2pin[i + 1]; ++s)
        forPartners(d, a, d.flatNode[s], minD, ignoreDeg, [&](int, int nq, float) {
            if (nq < d.nMov) ++c;
        });
    cnt[r] = c;
}

// Row r: entries (column = movable node, -w), FP64 diagonal and right-hand
// side (assemble, quadratic_placer.cpp:75-112), summed in the visiting order.
__global__ void k_qp_fill(QpDesign d, float minD, int ignoreDeg, const int64_t* rowStart,
                          int32_t* col, double* val, double* diag, double* rhs) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 2 * d.nMov) return;
    const int a = r >= d.nMov, i = r - a * d.nMov;
    const float* pos = d.pos[a];
    const float* off = d.off[a];
    int64_t e = rowStart[r];
    double dg = 0.0, rh = 0.0;
    for (int s = d.node2pin[i]; s < d.node2pin[i + 1]; ++s) {
        const int p = d.flatNode[s];
        forPartners(d, a, p, minD, ignoreDeg, [&](int q, int nq, float w) {
            const double wd = static_cast<double>(w);
            dg += wd;
            if (nq < d.nMov) {
                col[e] = nq;
                val[e] = -wd;
                ++e;
                rh += -wd * (static_cast<double>(off[p]) - static_cast<double>(off[q]));
            } else {
                rh += wd * (static_cast<double>(pos[nq] + off[q]) - static_cast<double>(off[p]));
            }
        });
    }
    if (dg == 0.0) {
        dg = 1e-6;
        rh += 1e-6 * static_cast<double>(pos[i]);
    }
    diag[r] = dg;
    rhs[r] = rh;
}

// ---- BiCGSTAB ---------------------------------------------------------------
__device__ inline double spmvRow(const QpVec& q, int a, int i, const double* in) {
    const int r = a * q.n + i;
    const double* base = in + static_cast<long>(a) * q.n;
    double acc = q.diag[r] * in[r];
    for (int64_t e = q.rowStart[r]; e < q.rowStart[r + 1]; ++e) acc += q.val[e] * base[q.col[e]];
    return acc;
}

// Block sum of K values (every thread passes its own), then the axis' last
// block to finish sums the block partials in block order and returns true
// (with the totals in out); all other blocks return false.
template <int K>
__device__ bool axisReduce(const QpVec& q, int a, double (&v)[K], double (&out)[K]) {
    __shared__ double sh[K][kQpBlock / 64];
    __shared__ bool last;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    double* part = q.part + static_cast<long>(a) * kQpMaxBlocks * 4;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int w = 0; w < kQpBlock / 64; ++w) s += sh[k][w];
            part[blockIdx.x * 4 + k] = s;
        }
        __threadfence();
        const unsigned prev = atomicAdd(&q.count[a], 1u);
        last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    // the other blocks' partials: fetched by the whole block at once, then
    // added in block order by one thread
    __shared__ double stage[K][kQpMaxBlocks];
    __threadfence();
    for (unsigned bk = threadIdx.x; bk < gridDim.x; bk += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; ++k) stage[k][bk] = part[bk * 4 + k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (unsigned bk = 0; bk < gridDim.x; ++bk) s += stage[k][bk];
            out[k] = s;
        }
        q.count[a] = 0;   // ready for the next reduction
    }
    return threadIdx.x == 0;
}

// r = b - A x, r0 = r, p = v = 0; scalars from b.b and r.r.
__global__ __launch_bounds__(kQpBlock) void k_bicg_init(QpVec q, double tol) {
    const int a = blockIdx.y;
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += gridDim.x * blockDim.x) {
        const int r = a * q.n + i;
        const double ri = q.b[r] - spmvRow(q, a, i, q.x);
        q.r[r] = ri;
        q.r0[r] = ri;
        q.p[r] = 0.0;
        q.v[r] = 0.0;
        acc[0] += q.b[r] * q.b[r];
        acc[1] += ri * ri;
    }
    double tot[2];
    if (axisReduce<2>(q, a, acc, tot)) {
        QpScal& s = q.sc[a];
        s.bnorm2 = tot[0];
        s.rr = tot[1];
        s.rho = tot[1];      // r0 . r
        s.rhoPrev = 1.0;
        s.alpha = 1.0;
        s.w = 1.0;
        s.iter = 0;
        s.tol2 = tol * tol * tot[0];
        s.done = tot[0] == 0.0 || !(tot[1] > s.tol2) || q.maxIter <= 0;
    }
}

// beta = (rho / rhoPrev)(alpha / w); p = r + beta (p - w v); y = p / diag.
__global__ __launch_bounds__(kQpBlock) void k_bicg_p(QpVec q) {
    const int a = blockIdx.y;
    QpScal s = q.sc[a];
    if (s.done) return;
    if (s.rhoPrev == 0.0 || s.w == 0.0) {   // breakdown: stop where it is
        if (blockIdx.x == 0 && threadIdx.x == 0) q.sc[a].done = 1;
        return;
    }
    const double beta = (s.rho / s.rhoPrev) * (s.alpha / s.w);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += gridDim.x * blockDim.x) {
        const int r = a * q.n + i;
        const double pi = q.r[r] + beta * (q.p[r] - s.w * q.v[r]);
        q.p[r] = pi;
        q.y[r] = pi / q.diag[r];
    }
}

// v = A y; alpha = rho / (r0 . v).
__global__ __launch_bounds__(kQpBlock) void k_bicg_v(QpVec q) {
    const int a = blockIdx.y;
    if (q.sc[a].done) return;
    double acc[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += gridDim.x * blockDim.x) {
        const int r = a * q.n + i;
        const double vi = spmvRow(q, a, i, q.y);
        q.v[r] = vi;
        acc[0] += q.r0[r] * vi;
    }
    double tot[1];
    if (axisReduce<1>(q, a, acc, tot)) {
        QpScal& s = q.sc[a];
        if (tot[0] == 0.0) s.done = 1;
        else s.alpha = s.rho / tot[0];
    }
}

// s = r - alpha v; z = s / diag.
__global__ __launch_bounds__(kQpBlock) void k_bicg_s(QpVec q) {
    const int a = blockIdx.y;
    const QpScal s = q.sc[a];
    if (s.done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += gridDim.x * blockDim.x) {
        const int r = a * q.n + i;
        const double si = q.r[r] - s.alpha * q.v[r];
        q.s[r] = si;
        q.z[r] = si / q.diag[r];
    }
}

// t = A z; w = (t . s) / (t . t).
__global__ __launch_bounds__(kQpBlock) void k_bicg_t(QpVec q) {
    const int a = blockIdx.y;
    if (q.sc[a].done) return;
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += gridDim.x * blockDim.x) {
        const int r = a * q.n + i;
        const double ti = spmvRow(q, a, i, q.z);
        q.t[r] = ti;
        acc[0] += ti * ti;
        acc[1] += ti * q.s[r];
    }
    double tot[2];
    if (axisReduce<2>(q, a, acc, tot)) q.sc[a].w = tot[0] > 0.0 ? tot[1] / tot[0] : 0.0;
}

// x += alpha y + w z; r = s - w t; the next rho and the stop test.
__global__ __launch_bounds__(kQpBlock) void k_bicg_x(QpVec q) {
    const int a = blockIdx.y;
    const QpScal s = q.sc[a];
    if (s.done) return;
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += gridDim.x * blockDim.x) {
        const int r = a * q.n + i;
        q.x[r] += s.alpha * q.y[r] + s.w * q.z[r];
        const double ri = q.s[r] - s.w * q.t[r];
        q.r[r] = ri;
        acc[0] += ri * ri;
        acc[1] += q.r0[r] * ri;
    }
    double tot[2];
    if (axisReduce<2>(q, a, acc, tot)) {
        QpScal& sc = q.sc[a];
        sc.rr = tot[0];
        sc.rhoPrev = sc.rho;
        sc.rho = tot[1];
        sc.iter += 1;
        sc.done = !(tot[0] > sc.tol2) || sc.iter >= q.maxIter;
    }
}

// Solution back to positions, clamped into the core (quadratic_placer.cpp:120-128);
// the clamped value is the next warm start.
__global__ void k_qp_clamp(QpDesign d, QpVec q) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 2 * d.nMov) return;
    const int a = r >= d.nMov, i = r - a * d.nMov;
    float v = q.sc[a].bnorm2 == 0.0 ? 0.f : static_cast<float>(q.x[r]);
    float* pos = d.pos[a];
    if (!isfinite(v)) v = pos[i];
    const float half = 0.5f * (a ? d.h[i] : d.w[i]);
    const float lo = d.core[a] + half, hi = d.core[a + 2] - half;
    v = lo <= hi ? fminf(fmaxf(v, lo), hi) : 0.5f * (lo + hi);
    pos[i] = v;
    q.x[r] = static_cast<double>(v);
}

// HPWL (computeHPWL, src/db/hpwl.cpp:13-39): per-block FP64 partials, summed
// in block order by the last block.
__global__ __launch_bounds__(kQpBlock) void k_qp_hpwl(QpDesign d, double* part, unsigned* count,
                                                      double* out) {
    double acc = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d.nNets; k += gridDim.x * blockDim.x) {
        const int b = d.net2pin[k], e = d.net2pin[k + 1];
        if (e - b < 2) continue;
        float lox = 3.402823466e38f, hix = -3.402823466e38f, loy = lox, hiy = hix;
        for (int t = b; t < e; ++t) {
            const int p = d.flatNet[t];
            const int n = d.pin2node[p];
            const float px = d.pos[0][n] + d.off[0][p], py = d.pos[1][n] + d.off[1][p];
            lox = fminf(lox, px);
            hix = fmaxf(hix, px);
            loy = fminf(loy, py);
            hiy = fmaxf(hiy, py);
        }
        acc += static_cast<double>(hix - lox) + static_cast<double>(hiy - loy);
    }
    __shared__ double sh[kQpBlock / 64];
    __shared__ bool last;
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kQpBlock / 64; ++w) s += sh[w];
        part[blockIdx.x] = s;
        __threadfence();
        last = atomicAdd(count, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __shared__ double stage[kQpMaxBlocks];
    __threadfence();
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) stage[b] = part[b];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (unsigned b = 0; b < gridDim.x; ++b) s += stage[b];
        *out = s;
        *count = 0;
    }
}

template <typename T>
T* qalloc(size_t n) {
    void* p = nullptr;
    SPHIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
T* qupload(const T* h, size_t n, hipStream_t s) {
    T* d = qalloc<T>(n);
    if (n) SPHIP_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

// Device copy of a design and the work buffers of the quadratic placer.
class QpSolver {
public:
    QpSolver(const sp_hip_design& d, const float* netW) {
        requireDevice();
        validateDesign(d);
        SPHIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        n_ = d.num_movable;
        QpDesign& q = dd_;
        q.nMov = d.num_movable;
        q.nNodes = d.num_nodes;
        q.nPins = d.num_pins;
        q.nNets = d.num_nets;
        const size_t nodes = d.num_nodes, pins = d.num_pins, nets = d.num_nets;
        q.w = track(qupload(d.node_w, nodes, s_));
        q.h = track(qupload(d.node_h, nodes, s_));
        q.pin2node = track(qupload(d.pin2node, pins, s_));
        q.pin2net = track(qupload(d.pin2net, pins, s_));
        q.off[0] = track(qupload(d.pin_offset_x, pins, s_));
        q.off[1] = track(qupload(d.pin_offset_y, pins, s_));
        q.net2pin = track(qupload(d.net2pin_start, nets + 1, s_));
        q.flatNet = track(qupload(d.flat_net2pin, pins, s_));
        q.node2pin = track(qupload(d.node2pin_start, nodes + 1, s_));
        q.flatNode = track(qupload(d.flat_node2pin, pins, s_));
        q.netW = netW ? track(qupload(netW, nets, s_)) : nullptr;
        q.pos[0] = track(qupload(d.node_x, nodes, s_));
        q.pos[1] = track(qupload(d.node_y, nodes, s_));
        q.bnd = track(qalloc<int32_t>(4 * nets));
        q.core[0] = d.core_lx;
        q.core[1] = d.core_ly;
        q.core[2] = d.core_hx;
        q.core[3] = d.core_hy;
        const size_t rows = 2 * static_cast<size_t>(n_);
        cnt_ = track(qalloc<int64_t>(rows + 1));
        rowStart_ = track(qalloc<int64_t>(rows + 1));
        for (double** v : {&v_.x, &v_.r, &v_.r0, &v_.p, &v_.v, &v_.y, &v_.s, &v_.z, &v_.t, &v_.b,
                           &v_.diag})
            *v = track(qalloc<double>(rows));
        v_.n = n_;
        v_.sc = track(qalloc<QpScal>(2));
        v_.part = track(qalloc<double>(2 * kQpMaxBlocks * 4));
        v_.count = track(qalloc<unsigned>(2));
        hpwlPart_ = track(qalloc<double>(kQpMaxBlocks));
        hpwlCount_ = track(qalloc<unsigned>(1));
        hpwl_ = track(qalloc<double>(1));
        SPHIP_CHECK(hipMemsetAsync(v_.count, 0, 2 * sizeof(unsigned), s_));
        SPHIP_CHECK(hipMemsetAsync(hpwlCount_, 0, sizeof(unsigned), s_));
        SPHIP_CHECK(hipMemsetAsync(v_.sc, 0, 2 * sizeof(QpScal), s_));
        SPHIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hScal_), 2 * sizeof(QpScal) + 16,
                                  hipHostMallocDefault));
        size_t tmp = 0;
        SPHIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt_, rowStart_,
                                                     static_cast<int>(rows + 1), s_));
        scanBytes_ = tmp;
        scanTmp_ = track(qalloc<unsigned char>(tmp));
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }
    ~QpSolver() {
        if (s_) (void)hipStreamSynchronize(s_);
        for (void* p : bufs_) (void)hipFree(p);
        if (col_) (void)hipFree(col_);
        if (val_) (void)hipFree(val_);
        if (hScal_) (void)hipHostFree(hScal_);
        if (s_) (void)hipStreamDestroy(s_);
    }

    // Bound pins and the B2B system of both axes at the current positions;
    // returns the entry count.
    int64_t assemble(float minDist, int ignoreDeg) {
        const float minD = std::max(1e-6f, minDist);
        const int rows = 2 * n_;
        if (dd_.nNets > 0)
            k_qp_bounds<<<ceilDiv(dd_.nNets, kQpBlock), kQpBlock, 0, s_>>>(dd_);
        SPHIP_CHECK(hipMemsetAsync(cnt_ + rows, 0, sizeof(int64_t), s_));
        if (rows > 0) k_qp_count<<<ceilDiv(rows, kQpBlock), kQpBlock, 0, s_>>>(dd_, minD, ignoreDeg, cnt_);
        SPHIP_CHECK(hipGetLastError());
        SPHIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scanTmp_, scanBytes_, cnt_, rowStart_, rows + 1, s_));
        int64_t nnz = 0;
        SPHIP_CHECK(hipMemcpyAsync(&nnz, rowStart_ + rows, sizeof(int64_t), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
        if (nnz > cap_) {
            if (col_) SPHIP_CHECK(hipFree(col_));
            if (val_) SPHIP_CHECK(hipFree(val_));
            col_ = nullptr;
            val_ = nullptr;
            cap_ = nnz + nnz / 4 + 1024;
            col_ = qalloc<int32_t>(cap_);
            val_ = qalloc<double>(cap_);
        }
        if (rows > 0)
            k_qp_fill<<<ceilDiv(rows, kQpBlock), kQpBlock, 0, s_>>>(dd_, minD, ignoreDeg, rowStart_,
                                                                   col_, val_, v_.diag, v_.b);
        SPHIP_CHECK(hipGetLastError());
        v_.rowStart = rowStart_;
        v_.col = col_;
        v_.val = val_;
        return nnz;
    }

    // BiCGSTAB on both axes from the warm start in v_.x; errors per axis.
    void solve(double tol, int maxIter, double err[2]) {
        v_.maxIter = maxIter;
        const dim3 grid(std::max(1, std::min(kQpMaxBlocks, ceilDiv(n_, kQpBlock))), 2);
        k_bicg_init<<<grid, kQpBlock, 0, s_>>>(v_, tol);
        for (int it = 0; it < maxIter; ++it) {
            k_bicg_p<<<grid, kQpBlock, 0, s_>>>(v_);
            k_bicg_v<<<grid, kQpBlock, 0, s_>>>(v_);
            k_bicg_s<<<grid, kQpBlock, 0, s_>>>(v_);
            k_bicg_t<<<grid, kQpBlock, 0, s_>>>(v_);
            k_bicg_x<<<grid, kQpBlock, 0, s_>>>(v_);
        }
        SPHIP_CHECK(hipGetLastError());
        SPHIP_CHECK(hipMemcpyAsync(hScal_, v_.sc, 2 * sizeof(QpScal), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
        for (int a = 0; a < 2; ++a) {
            const QpScal& s = hScal_[a];
            const double e = s.bnorm2 > 0.0 ? std::sqrt(s.rr / s.bnorm2) : 0.0;
            err[a] = std::isfinite(e) ? e : 0.0;
        }
    }

    void clampAndHpwl(double* hpwl) {
        const int rows = 2 * n_;
        if (rows > 0) k_qp_clamp<<<ceilDiv(rows, kQpBlock), kQpBlock, 0, s_>>>(dd_, v_);
        const int blocks = std::max(1, std::min(kQpMaxBlocks, ceilDiv(dd_.nNets, kQpBlock)));
        k_qp_hpwl<<<blocks, kQpBlock, 0, s_>>>(dd_, hpwlPart_, hpwlCount_, hpwl_);
        SPHIP_CHECK(hipGetLastError());
        SPHIP_CHECK(hipMemcpyAsync(hpwl, hpwl_, sizeof(double), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }

    // Movable nodes to the core centre; the warm start follows.
    void startAtCentre() {
        std::vector<double> xs(2 * static_cast<size_t>(n_));
        const float cx = 0.5f * (dd_.core[0] + dd_.core[2]), cy = 0.5f * (dd_.core[1] + dd_.core[3]);
        std::vector<float> px(n_, cx), py(n_, cy);
        for (int i = 0; i < n_; ++i) {
            xs[i] = cx;
            xs[n_ + i] = cy;
        }
        if (n_ > 0) {
            SPHIP_CHECK(hipMemcpyAsync(dd_.pos[0], px.data(), n_ * sizeof(float), hipMemcpyHostToDevice, s_));
            SPHIP_CHECK(hipMemcpyAsync(dd_.pos[1], py.data(), n_ * sizeof(float), hipMemcpyHostToDevice, s_));
            SPHIP_CHECK(hipMemcpyAsync(v_.x, xs.data(), xs.size() * sizeof(double), hipMemcpyHostToDevice, s_));
        }
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }

    void positions(float* x, float* y) {
        SPHIP_CHECK(hipMemcpyAsync(x, dd_.pos[0], dd_.nNodes * sizeof(float), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipMemcpyAsync(y, dd_.pos[1], dd_.nNodes * sizeof(float), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }

    // One axis of the assembled system to the host (rows of that axis).
    void system(int axis, double* diag, double* rhs, int64_t* rowStart, int32_t* col, double* val) {
        const size_t off = static_cast<size_t>(axis) * n_;
        std::vector<int64_t> rs(n_ + 1);
        SPHIP_CHECK(hipMemcpyAsync(rs.data(), rowStart_ + off, (n_ + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
        const int64_t base = rs[0], nnz = rs[n_] - rs[0];
        for (int i = 0; i <= n_; ++i) rowStart[i] = rs[i] - base;
        SPHIP_CHECK(hipMemcpyAsync(diag, v_.diag + off, n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipMemcpyAsync(rhs, v_.b + off, n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
        if (nnz > 0) {
            SPHIP_CHECK(hipMemcpyAsync(col, col_ + base, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s_));
            SPHIP_CHECK(hipMemcpyAsync(val, val_ + base, nnz * sizeof(double), hipMemcpyDeviceToHost, s_));
        }
        SPHIP_CHECK(hipStreamSynchronize(s_));
    }
    int64_t axisEntries(int axis) {
        int64_t ends[2];
        SPHIP_CHECK(hipMemcpyAsync(&ends[0], rowStart_ + static_cast<size_t>(axis) * n_, sizeof(int64_t), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipMemcpyAsync(&ends[1], rowStart_ + static_cast<size_t>(axis + 1) * n_, sizeof(int64_t), hipMemcpyDeviceToHost, s_));
        SPHIP_CHECK(hipStreamSynchronize(s_));
        return ends[1] - ends[0];
    }

    int movable() const { return n_; }

private:
    template <typename T>
    T* track(T* p) {
        bufs_.push_back(const_cast<void*>(static_cast<const void*>(p)));
        return p;
    }

    hipStream_t s_ = nullptr;
    int n_ = 0;
    QpDesign dd_{};
    QpVec v_{};
    std::vector<void*> bufs_;
    int64_t *cnt_ = nullptr, *rowStart_ = nullptr;
    int32_t* col_ = nullptr;
    double* val_ = nullptr;
    int64_t cap_ = 0;
    void* scanTmp_ = nullptr;
    size_t scanBytes_ = 0;
    double *hpwlPart_ = nullptr, *hpwl_ = nullptr;
    unsigned* hpwlCount_ = nullptr;
    QpScal* hScal_ = nullptr;
};

void checkQpConfig(const sp_qp_config& c) {
    if (c.max_iter < 1 || c.solver_max_iter < 0 || !(c.tol >= 0.f) || !(c.outer_tol >= 0.f) ||
        !(c.min_distance >= 0.f) || c.ignore_net_degree < 2)
        throw InvalidArg("quadratic placer: bad configuration");
}

}  // namespace
}  // namespace sphip

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_qp_place(const sp_hip_design* d, const float* net_weight, const sp_qp_config* cfg,
                float* x, float* y, sp_hip_iter_metrics* rows, int max_rows,
                int32_t* iterations) {
    return guarded([&] {
        if (!d || !cfg || !x || !y) throw InvalidArg("null argument");
        sphip::checkQpConfig(*cfg);
        sphip::QpSolver q(*d, net_weight);
        if (iterations) *iterations = 0;
        if (q.movable() == 0) {
            q.positions(x, y);
            return;
        }
        q.startAtCentre();
        double last = 0.0;
        int iter = 0;
        for (;; ++iter) {
            const auto t0 = std::chrono::steady_clock::now();
            q.assemble(cfg->min_distance, cfg->ignore_net_degree);
            double err[2];
            q.solve(cfg->tol, cfg->solver_max_iter, err);
            double hp = 0.0;
            q.clampAndHpwl(&hp);
            if (rows && iter < max_rows) {
                sp_hip_iter_metrics m{};
                m.iter = iter;
                m.hpwl = hp;
                m.grad_norm_wl = std::max(err[0], err[1]);
                m.elapsed_ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                rows[iter] = m;
            }
            const bool stalled = iter > 4 && last > 0.0 &&
                                 std::fabs(hp - last) < static_cast<double>(cfg->outer_tol) * last;
            last = hp;
            if (stalled || iter + 1 >= cfg->max_iter) break;
        }
        if (iterations) *iterations = iter + 1;
        q.positions(x, y);
    });
}

int sp_qp_b2b_system(const sp_hip_design* d, const float* net_weight, int axis,
                     float min_distance, int32_t ignore_net_degree, double* diag, double* rhs,
                     int64_t* row_start, int32_t* col, double* val, int64_t cap, int64_t* nnz) {
    return guarded([&] {
        if (!d || !nnz || (axis != 0 && axis != 1)) throw InvalidArg("bad argument");
        if (ignore_net_degree < 2) throw InvalidArg("ignore_net_degree must be >= 2");
        sphip::QpSolver q(*d, net_weight);
        q.assemble(min_distance, ignore_net_degree);
        *nnz = q.axisEntries(axis);
        if (!diag) return;
        if (!rhs || !row_start || (*nnz > 0 && (!col || !val))) throw InvalidArg("null output");
        if (cap < *nnz) throw InvalidArg("entry buffer too small");
        q.system(axis, diag, rhs, row_start, col, val);
    });
}

}  // extern "C"
