// HIP status checking for host code of the LMM engine.
#pragma once

#include <hip/hip_runtime.h>
#include <string>
#include "fm_internal.hpp"

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t fm_hip_status_ = (expr);                                                      \
        if (fm_hip_status_ != hipSuccess) {                                                      \
            const int fm_code_ = (fm_hip_status_ == hipErrorOutOfMemory) ? FM_ERR_OUT_OF_MEMORY   \
                                                                         : FM_ERR_DEVICE;         \
            throw ::fm::FmError(fm_code_, std::string(#expr) + " failed: " +                     \
                                          hipGetErrorString(fm_hip_status_));                    \
        }                                                                                        \
    } while (0)
