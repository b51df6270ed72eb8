// Error handling shared by the C-ABI entry points: exceptions thrown inside
// an entry point become an sp_hip_status and a thread-local message.
#pragma once

#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/superplace_hip.h"
#include "common.h"

namespace sphip {

inline thread_local std::string g_lastError;

struct InvalidArg : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NonFinite : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <typename F>
inline int guarded(F&& f) {
    try {
        f();
        return SP_HIP_OK;
    } catch (const InvalidArg& e) {
        g_lastError = e.what();
        return SP_HIP_EINVAL;
    } catch (const std::invalid_argument& e) {
        g_lastError = e.what();
        return SP_HIP_EINVAL;
    } catch (const NonFinite& e) {
        g_lastError = e.what();
        return SP_HIP_ENONFINITE;
    } catch (const NoDevice& e) {
        g_lastError = e.what();
        return SP_HIP_ENODEV;
    } catch (const HipError& e) {
        g_lastError = e.what();
        return std::strstr(e.what(), "out of memory") ? SP_HIP_EOOM : SP_HIP_EDEVICE;
    } catch (const std::bad_alloc& e) {
        g_lastError = "host allocation failed";
        return SP_HIP_EOOM;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return SP_HIP_EINTERNAL;
    }
}

// Throws NoDevice when no HIP device is visible.
// Throws NoDevice without a HIP device.  A device seen once stays (the
// count is taken again only until the first success).
inline void requireDevice() {
    static std::atomic<bool> seen{false};
    if (seen.load(std::memory_order_relaxed)) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        throw NoDevice("no HIP device available");
    }
    seen.store(true, std::memory_order_relaxed);
}

// Checks the counts, CSR spans and index ranges of a design passed in.
inline void validateDesign(const sp_hip_design& d) {
    if (d.num_movable < 0 || d.num_nodes < d.num_movable || d.num_fillers < 0 ||
        d.num_pins < 0 || d.num_nets < 0 || d.num_rows < 0)
        throw InvalidArg("design: inconsistent counts");
    const int total = d.num_nodes + d.num_fillers;
    if (total > 0 && (!d.node_x || !d.node_y || !d.node_w || !d.node_h || !d.node_flags))
        throw InvalidArg("design: missing node arrays");
    if (!d.net2pin_start || !d.node2pin_start)
        throw InvalidArg("design: missing CSR starts");
    if (d.num_pins > 0 && (!d.pin_offset_x || !d.pin_offset_y || !d.pin2node || !d.pin2net ||
                           !d.flat_net2pin || !d.flat_node2pin))
        throw InvalidArg("design: missing pin arrays");
    if (d.net2pin_start[0] != 0 || d.net2pin_start[d.num_nets] != d.num_pins)
        throw InvalidArg("design: net2pin_start does not span the pins");
    if (d.node2pin_start[0] != 0 || d.node2pin_start[total] != d.num_pins)
        throw InvalidArg("design: node2pin_start does not span the pins");
    for (int k = 0; k < d.num_nets; ++k)
        if (d.net2pin_start[k + 1] < d.net2pin_start[k]) throw InvalidArg("design: net CSR not monotone");
    for (int i = 0; i < total; ++i)
        if (d.node2pin_start[i + 1] < d.node2pin_start[i])
            throw InvalidArg("design: node CSR not monotone");
    for (int t = 0; t < d.num_pins; ++t) {
        if (d.flat_net2pin[t] < 0 || d.flat_net2pin[t] >= d.num_pins ||
            d.flat_node2pin[t] < 0 || d.flat_node2pin[t] >= d.num_pins)
            throw InvalidArg("design: CSR pin index out of range");
        if (d.pin2node[t] < 0 || d.pin2node[t] >= d.num_nodes)
            throw InvalidArg("design: pin2node out of range");
        if (d.pin2net[t] < 0 || d.pin2net[t] >= d.num_nets)
            throw InvalidArg("design: pin2net out of range");
    }
    if (!(d.core_hx > d.core_lx) || !(d.core_hy > d.core_ly)) throw InvalidArg("design: empty core");
    for (int i = 0; i < total; ++i)
        if (!std::isfinite(d.node_x[i]) || !std::isfinite(d.node_y[i]))
            throw NonFinite("design: non-finite node coordinate");
}

}  // namespace sphip
