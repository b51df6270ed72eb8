// Shared definitions for the gfx950 kernels and their host launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>
#include <stdexcept>
#include <string>

namespace sphip {

#define SPHIP_CHECK(expr)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::sphip::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

constexpr int kBlock = 256;   // threads per block for the elementwise/gather kernels

// MI355X deals a launch's workgroups to its 8 XCDs in turn (workgroup b to
// XCD b % 8), and each XCD has its own L2.  Logical block
// (b % 8) (n / 8) + b / 8 gives XCD x the logical blocks [x n/8, (x+1) n/8):
// a contiguous share of whatever the blocks index.  Bijective on [0, n)
// when 8 divides n; the identity otherwise.
constexpr int kXcds = 8;
__host__ __device__ inline int xcdBlock(int b, int n) {
    return n % kXcds == 0 ? (b % kXcds) * (n / kXcds) + b / kXcds : b;
}

// Device bytes allocated by this thread through the library's allocation
// helpers: an instance's footprint is the difference across its set-up.
inline thread_local size_t g_devBytes = 0;

inline int ceilDiv(long a, long b) { return static_cast<int>((a + b - 1) / b); }
inline int roundUp(int a, int b) { return (a + b - 1) / b * b; }

// Node flags (src/db/place_db.h:17-22).
enum : uint8_t { F_MACRO = 1, F_FIXED = 2, F_FILLER = 4, F_NI = 8 };

// Per-evaluation FP64 scalars, written by the last block of each reduction.
struct EvalScalars {
    double tau;        // overflow, bin_grid.cpp:373-393
    double wl;         // WA total, wa_wirelength.cpp:119-173
    double hpwl;       // hpwl.cpp:13-39
    double sum_wl;     // sum |dWL|, eplace.cpp:168-185
    double sum_den;    // sum |F|,  eplace.cpp:186-190
    double dx2;        // ||u - u_prev||^2
    double dg2;        // ||g - g_prev||^2
    double sy;         // (u - u_prev).(g - g_prev)
    double gg;         // ||g||^2
    double nonfinite;  // count of non-finite gradient entries
};
constexpr int kNumScalars = sizeof(EvalScalars) / sizeof(double);

// Grid geometry shared by scatter, finalize and gather.
template <typename T>
struct GridGeom {
    int dim;        // bins per side
    int ld;         // row stride of the padded field / rho layout
    T stepX, stepY, originX, originY;
    T target;       // target density (charge weight of macros)
};

// Optional event pairs timing the kernels of one family: while set, the
// k-th kernel launched records start[k] as its dispatch begins and stop[k]
// as it ends (hipExtLaunchKernel: the timestamps of the dispatch itself,
// without the gaps between launches).  Only for launches outside graph
// capture; at most kMaxTimed kernels per family.
constexpr int kMaxTimed = 4;
struct LaunchTimer {
    hipEvent_t start[kMaxTimed] = {}, stop[kMaxTimed] = {};
    int used = 0;
};
inline thread_local LaunchTimer* tlsTimer = nullptr;

template <typename... Params, typename... Args>
inline void launchKernel(void (*kernel)(Params...), dim3 grid, dim3 block, uint32_t shm,
                         hipStream_t s, Args... args) {
    LaunchTimer* t = tlsTimer;
    if (t && t->used < kMaxTimed) {
        hipExtLaunchKernelGGL(kernel, grid, block, shm, s, t->start[t->used], t->stop[t->used], 0,
                              static_cast<Params>(args)...);
        ++t->used;
    } else {
        kernel<<<grid, block, shm, s>>>(static_cast<Params>(args)...);
    }
}

// Raises a kernel's dynamic LDS limit to at least `bytes`.  The attribute
// belongs to the kernel, not to a launch, so instances with different
// launch shapes only ever raise it: lowering it for a small grid would
// break a larger grid's launches of the same kernel in another instance.
inline void raiseLdsLimit(const void* kernel, size_t bytes) {
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> limits;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& kv : limits)
        if (kv.first == kernel) {
            if (bytes <= kv.second) return;
            SPHIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(bytes)));
            kv.second = bytes;
            return;
        }
    SPHIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(bytes)));
    limits.emplace_back(kernel, bytes);
}

// Pairs stored interleaved (x0, y0, x1, y1, ...) and moved as one vector access.
template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };
template <typename T>
__host__ __device__ inline typename Vec2<T>::type load2(const T* p, long i) {
    return reinterpret_cast<const typename Vec2<T>::type*>(p)[i];
}
template <typename T>
__host__ __device__ inline void store2(T* p, long i, T a, T b) {
    typename Vec2<T>::type v;
    v.x = a;
    v.y = b;
    reinterpret_cast<typename Vec2<T>::type*>(p)[i] = v;
}

}  // namespace sphip
