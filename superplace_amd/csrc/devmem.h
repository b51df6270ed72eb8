// Memory of a placement instance, arranged so that setting an instance up
// and tearing it down take a handful of runtime calls:
//  - DevArena: all of an instance's device arrays in one allocation, laid
//    out while the host side is built and placed at once: one hipMalloc,
//    one host-to-device copy of every array that has host contents (packed
//    at the front) and one zero fill of the rest.  A separate allocation and
//    pageable copy per array cost 30-90 microseconds each, about 60 of them
//    per instance.
//  - Streams and small pinned host blocks come from per-process pools and
//    are reused by later instances: creating and destroying a stream costs
//    about a millisecond, a pinned allocation about 0.1 ms.  Pooled streams
//    and pinned chunks live until the process ends.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "common.h"

namespace sphip {

class DevArena {
public:
    DevArena() = default;
    DevArena(const DevArena&) = delete;
    DevArena& operator=(const DevArena&) = delete;
    ~DevArena() { release(); }

    // *dst receives the address of `count` zero-filled T at place().
    template <typename P>
    void zeros(P** dst, size_t count) {
        zero_.push_back({reinterpret_cast<void**>(const_cast<std::remove_const_t<P>**>(dst)),
                         zeroBytes_});
        zeroBytes_ += align(std::max<size_t>(count, 1) * sizeof(P));
    }

    // *dst receives the address of a device copy of h[0..count) converted
    // to the pointee type at place().  The converted copy is held here until
    // then.
    template <typename P, typename S>
    void upload(P** dst, const S* h, size_t count) {
        using T = std::remove_const_t<P>;
        const size_t off = image_.size();
        image_.resize(off + align(std::max<size_t>(count, 1) * sizeof(T)), 0);
        T* t = reinterpret_cast<T*>(image_.data() + off);
        for (size_t i = 0; i < count; ++i) t[i] = static_cast<T>(h[i]);
        up_.push_back({reinterpret_cast<void**>(const_cast<T**>(dst)), off});
    }
    template <typename P, typename S>
    void upload(P** dst, const std::vector<S>& h) {
        upload(dst, h.data(), h.size());
    }

    // Allocates the arena, sets every pointer reserved so far, copies the
    // host contents and zero-fills the rest on stream s, and waits for both.
    void place(hipStream_t s) {
        if (base_) throw std::logic_error("device arena placed twice");
        const size_t upBytes = image_.size();
        bytes_ = upBytes + zeroBytes_;
        SPHIP_CHECK(hipMalloc(&base_, std::max<size_t>(bytes_, 256)));
        g_devBytes += bytes_;
        char* b = static_cast<char*>(base_);
        for (const Slot& r : up_) *r.dst = b + r.off;
        for (const Slot& r : zero_) *r.dst = b + upBytes + r.off;
        if (upBytes)
            SPHIP_CHECK(hipMemcpyAsync(b, image_.data(), upBytes, hipMemcpyHostToDevice, s));
        if (zeroBytes_) SPHIP_CHECK(hipMemsetAsync(b + upBytes, 0, zeroBytes_, s));
        SPHIP_CHECK(hipStreamSynchronize(s));
        std::vector<unsigned char>().swap(image_);
        up_.clear();
        zero_.clear();
    }

    // Frees the arena (the caller has synchronised the streams using it).
    void release() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }

    size_t bytes() const { return bytes_; }

private:
    static size_t align(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
    struct Slot {
        void** dst;
        size_t off;
    };
    std::vector<unsigned char> image_;
    std::vector<Slot> up_, zero_;
    size_t zeroBytes_ = 0, bytes_ = 0;
    void* base_ = nullptr;
};

namespace pools {

inline std::mutex& mutex() {
    static std::mutex m;
    return m;
}

struct PooledStream {
    int device;
    hipStream_t stream;
};

inline std::vector<PooledStream>& streams() {
    static std::vector<PooledStream> v;
    return v;
}

constexpr size_t kPinnedBlock = 1024;      // bytes per pinned block
constexpr size_t kPinnedPerChunk = 64;     // blocks per pinned allocation

inline std::vector<void*>& pinnedFree() {
    static std::vector<void*> v;
    return v;
}

}  // namespace pools

// A non-blocking stream of device `device` (the current device), from the
// pool or newly created.
inline hipStream_t takeStream(int device) {
    {
        std::lock_guard<std::mutex> lock(pools::mutex());
        auto& v = pools::streams();
        for (size_t i = v.size(); i-- > 0;)
            if (v[i].device == device) {
                const hipStream_t s = v[i].stream;
                v.erase(v.begin() + static_cast<long>(i));
                return s;
            }
    }
    hipStream_t s = nullptr;
    SPHIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

// Returns an idle stream of `device` to the pool.
inline void giveStream(int device, hipStream_t s) {
    if (!s) return;
    std::lock_guard<std::mutex> lock(pools::mutex());
    pools::streams().push_back({device, s});
}

// A pinned host block of pools::kPinnedBlock bytes.
inline void* takePinned() {
    std::lock_guard<std::mutex> lock(pools::mutex());
    auto& v = pools::pinnedFree();
    if (v.empty()) {
        void* chunk = nullptr;
        SPHIP_CHECK(hipHostMalloc(&chunk, pools::kPinnedBlock * pools::kPinnedPerChunk,
                                  hipHostMallocDefault));
        for (size_t i = pools::kPinnedPerChunk; i-- > 0;)
            v.push_back(static_cast<char*>(chunk) + i * pools::kPinnedBlock);
    }
    void* p = v.back();
    v.pop_back();
    return p;
}

inline void givePinned(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(pools::mutex());
    pools::pinnedFree().push_back(p);
}

}  // namespace sphip
