// runtime_util.hpp -- error plumbing shared by the C-ABI translation units: HIP status checks
// that throw, and the wrapper that turns exceptions into status codes at the boundary.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/exaca_ca.h"

namespace exaca {

extern thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return EXACA_OK;
    }
    catch (const DeviceError &e) {
        g_last_error = e.what();
        return EXACA_ERR_DEVICE;
    }
    catch (const StateError &e) {
        g_last_error = e.what();
        return EXACA_ERR_STATE;
    }
    catch (const std::exception &e) {
        g_last_error = e.what();
        return EXACA_ERR_INPUT;
    }
    catch (...) {
        g_last_error = "unknown error";
        return EXACA_ERR_INPUT;
    }
}

}  // namespace exaca

#define HIP_OK(x) ::exaca::hip_check((x), #x)
