// Host-side helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/alsvinn_hip.h"
#include "avh_stage_inst.hpp"

struct avh_ctx {
    int device;
};

namespace avh {

inline std::string& last_error_slot() {
    static thread_local std::string s;
    return s;
}

inline avh_status fail(avh_status st, const std::string& msg) {
    last_error_slot() = msg;
    return st;
}

#define AVH_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::avh::fail(AVH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

inline int nvars_of(int eq) {
    switch (eq) {
    case AVH_EQ_BURGERS: return 1;
    case AVH_EQ_EULER1: return 3;
    case AVH_EQ_EULER2: return 4;
    case AVH_EQ_EULER3: return 5;
    }
    return 0;
}

inline int ghosts_of(int recon) {
    switch (recon) {
    case AVH_RECON_NONE: return 1;
    case AVH_RECON_WENO2: return 2;
    case AVH_RECON_ENO2: return 2;
    case AVH_RECON_ENO3: return 3;
    }
    return 0;
}

inline Launchers launchers_for(int equation, int recon, int flux, int dim) {
    switch (equation) {
    case AVH_EQ_BURGERS: return get_launchers_burgers(recon, flux, dim);
    case AVH_EQ_EULER1: return get_launchers_euler1(recon, flux, dim);
    case AVH_EQ_EULER2: return get_launchers_euler2(recon, flux, dim);
    case AVH_EQ_EULER3: return get_launchers_euler3(recon, flux, dim);
    }
    return Launchers();
}

}  // namespace avh
