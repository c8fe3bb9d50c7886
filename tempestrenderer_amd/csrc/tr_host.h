/*
 * tr_host.h - host-side helpers shared by the C ABI translation units:
 * the thread-local error message, HIP error propagation and a growable
 * device buffer.
 */
#ifndef TR_HOST_H
#define TR_HOST_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/tempest_rt.h"

namespace trk {

extern thread_local std::string g_error;

inline tr_status fail(tr_status st, const std::string& msg)
{
    g_error = msg;
    return st;
}

#define TR_HIP(call)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return ::trk::fail(TR_ERROR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T> struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    hipError_t reserve(size_t n)
    {
        if (n <= count && ptr) return hipSuccess;
        release();
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc((void**)&ptr, bytes);
        if (e == hipSuccess) count = n;
        return e;
    }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace trk

#endif
