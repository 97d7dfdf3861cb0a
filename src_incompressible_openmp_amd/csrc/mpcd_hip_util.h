/*
 * mpcd_hip_util.h -- host-side helpers shared by the library's HIP
 * translation units: HIP error checks, device allocation, and the mapping
 * of exceptions to the C ABI's status codes and last-error message.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "mpcd_gpu.h"
#include "mpcd_host.h"

namespace mpcdk {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::mpcdk::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));               \
    } while (0)

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc(&p, n * sizeof(T)));
    return (T*)p;
}

/* Device memory owned by a scope: freed when the scope ends, also when an
 * error unwinds it.  Empty until given an allocation. */
template <class T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t n) : p(dalloc<T>(n)) {}
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

/* the calling thread's last error message (mpcd_gpu_last_error) */
void set_last_error(const std::string& m);

/* runs f, mapping its exceptions to a status code and the last error */
template <class F>
int guarded(F&& f) {
    try {
        f();
        return MPCD_OK;
    } catch (const ApiError& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const HipError& e) {
        set_last_error(e.what());
        return MPCD_E_HIP;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MPCD_E_INVALID;
    }
}

/* the closure surrogate as the simulation uses it (mpcd_net.hip): forward of
 * one grid of features [cell][4] on `st`; returns the device output
 * [cell][2], valid until the network's next call */
const float* net_forward_device(mpcd_gpu_net* net, const float* featDev, hipStream_t st);
void net_grid(const mpcd_gpu_net* net, int* Nx, int* Ny, int* perx, int* pery);

}  // namespace mpcdk
