// Shared definitions for the gfx950 kernels and their host launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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

}  // namespace sphip
