// HIP plumbing shared by the host-side launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cupbr.h"

namespace cupbr {

int set_error(int code, const std::string& msg, bool print = false);

#define CUPBR_HIP_CHECK(expr)                                                                              \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return ::cupbr::set_error(CUPBR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_), false); \
    } while (0)

// number of compute units of the current device (cached per device)
int device_cu_count();

// RAII device buffer that only grows
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t n) {
        if (n <= bytes) return CUPBR_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (n == 0) return CUPBR_OK;
        if (hipMalloc(&ptr, n) != hipSuccess) return set_error(CUPBR_ERR_OUT_OF_MEMORY, "hipMalloc failed", false);
        bytes = n;
        return CUPBR_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(ptr); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace cupbr
