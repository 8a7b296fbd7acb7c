// runtime_util.hpp -- error plumbing shared by the C-ABI translation units: HIP status checks
// that throw, and the wrapper that turns exceptions into status codes at the boundary.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../include/exaca_ca.h"
#include "ca_common.hpp"

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

// Device error flags (StepCtl::error) and their meaning, for the messages of the ABI calls.
inline std::string device_error_text(int flags) {
    std::string s = "device error flag " + std::to_string(flags) + " (";
    const char *sep = "";
    if (flags & 1) s += sep, s += "1: a cell list overflowed", sep = ", ";
    if (flags & 2) s += sep, s += "2: an active cell with GrainID 0", sep = ", ";
    if (flags & 8) s += sep, s += "8: a list-rebuild look-back did not end within its bound", sep = ", ";
    if (flags & ~11) s += sep, s += "unknown bits";
    return s + ")";
}

// Candidates staged in LDS per apply workgroup: kCandStage, or fewer when EXACA_CAND_STAGE
// (1 .. kCandStage) asks for it, so that tests can drive the apply pass's overflow list on small
// domains.
// Workgroups of the frontier kernels of a plain-path handle: EXACA_BLOCKS (8 .. 2048, a multiple
// of 8) overrides the size-based choice. Tests narrow it so that slices hold many claimants and
// the kernels' multi-round paths run.
inline int frontier_blocks_override() {
    const char *env = std::getenv("EXACA_BLOCKS");
    if (!env) return 0;
    const long v = std::strtol(env, nullptr, 10);
    if (v < 8 || v > 2048 || v % 8)
        throw std::invalid_argument("EXACA_BLOCKS must be a multiple of 8 in 8 .. 2048");
    return static_cast<int>(v);
}
inline uint32_t cand_stage_cap() {
    const char *env = std::getenv("EXACA_CAND_STAGE");
    if (!env) return kCandStage;
    const long v = std::strtol(env, nullptr, 10);
    if (v < 1 || v > static_cast<long>(kCandStage))
        throw std::invalid_argument("EXACA_CAND_STAGE must be 1 .. " + std::to_string(kCandStage));
    return static_cast<uint32_t>(v);
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
