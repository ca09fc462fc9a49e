// Host helpers shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include <string>

#include "../../include/lqt.h"
#include "internal.hpp"

namespace lqt {

lqt_status set_error(lqt_status code, const char* fmt, ...);

#define LQT_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::lqt::set_error(LQT_ERR_HIP, "%s failed: %s (%s:%d)", #call,                \
                                    hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define LQT_CHECK(cond, code, ...)                                                              \
    do {                                                                                        \
        if (!(cond)) return ::lqt::set_error(code, __VA_ARGS__);                                \
    } while (0)

// Per-class constants in double, following the operation order of
// TPORd.cc:625-645 (PCa) and :693-733 (GKs) / GPU3D.cu:858-909.
Consts<double> derive_consts(const lqt_params& p);
Consts<float> to_float(const Consts<double>& c);

// Upload helper.
template <class T>
lqt_status upload(T** dptr, const T* host, size_t count) {
    LQT_HIP(hipMalloc((void**)dptr, sizeof(T) * (count ? count : 1)));
    if (host && count) LQT_HIP(hipMemcpy(*dptr, host, sizeof(T) * count, hipMemcpyHostToDevice));
    return LQT_OK;
}

}  // namespace lqt
