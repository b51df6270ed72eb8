// The Poisson solve for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): sp_hip_poisson_solve_noalloc.
//
// The intermediates, tables and spectral scales of the FFT passes live in
// one file-scope device array, `scratch`, of kScratchFloats floats (enough
// for a 1024 x 1024 grid: 12.6 MB), which every GPU the library runs on
// holds until the process ends.  Host code never passes the array's name to
// a kernel (in host code it is only a stand-in symbol): the device address
// is looked up once per device with hipGetSymbolAddress, with that device
// current, and kept per device.  Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU (two streams, two threads) race;
// each call rewrites all of the array it uses before reading it.
#include <mutex>

#include "../../include/superplace_hip.h"
#include "abi.h"
#include "poisson.h"

namespace sphip {
namespace {

constexpr int kMaxSide = 1024;
constexpr size_t kScratchFloats =
    3 * static_cast<size_t>(kMaxSide) * kMaxSide + 10 * static_cast<size_t>(kMaxSide);
constexpr int kMaxDevices = 64;

__device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_scratchMu;
float* g_scratchDev[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the
// first call on that device).
float* scratchOnCurrentDevice() {
    int dev = -1;
    SPHIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw InvalidArg("no-allocation solve: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchDev[dev]) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p) {
            (void)hipGetLastError();
            throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        }
        g_scratchDev[dev] = static_cast<float*>(p);
    }
    return g_scratchDev[dev];
}

}  // namespace
}  // namespace sphip

extern "C" int sp_hip_poisson_solve_noalloc(int nx, int ny, float step_x, float step_y,
                                            const float* density_dev, float* out3_dev,
                                            void* stream) {
    return sphip::guarded([&] {
        sphip::requireDevice();
        auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
        if (!pow2(nx) || !pow2(ny) || nx < 64 || ny < 64 || nx > sphip::kMaxSide ||
            ny > sphip::kMaxSide)
            throw sphip::InvalidArg(
                "no-allocation solve: grid sides must be powers of two from 64 to 1024 (got " +
                std::to_string(nx) + " x " + std::to_string(ny) + ")");
        if (!(step_x > 0.f) || !(step_y > 0.f))
            throw sphip::InvalidArg("no-allocation solve: bin step must be positive");
        if (!density_dev || !out3_dev) throw sphip::InvalidArg("no-allocation solve: null pointer");
        if (sphip::fftSolveWorkFloats(nx, ny) > sphip::kScratchFloats)
            throw sphip::InvalidArg("no-allocation solve: grid exceeds the scratch array");
        float* work = sphip::scratchOnCurrentDevice();
        sphip::fftSolveInto(nx, ny, step_x, step_y, density_dev, out3_dev, work,
                            static_cast<hipStream_t>(stream));
    });
}
