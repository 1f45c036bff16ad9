// rc_host.h -- host-side helpers shared by the engine's translation units:
// the library's error slot and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace rc {

// Records the first error since the last clear; rc_last_error() returns it.
void set_error(const std::string& e);

#define HIP_OK(expr)                                                                          \
    do {                                                                                      \
        hipError_t err_ = (expr);                                                             \
        if (err_ != hipSuccess) {                                                             \
            ::rc::set_error(std::string("HIP error: ") + hipGetErrorString(err_) + " at " #expr); \
            return false;                                                                     \
        }                                                                                     \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool alloc(size_t count) {
        if (count <= n && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc failed");
            return false;
        }
        n = count;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace rc
