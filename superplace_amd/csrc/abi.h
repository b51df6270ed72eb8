// Error handling shared by the C-ABI entry points: exceptions thrown inside
// an entry point become an sp_hip_status and a thread-local message.
#pragma once

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

}  // namespace sphip
