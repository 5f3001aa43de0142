// Host helpers shared by the C-ABI implementations: error type, HIP status check, device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "spg_vbd.h"

namespace spgvbd {

struct VbdError : std::runtime_error {
    int code;
    VbdError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hipCheck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw VbdError(SPG_VBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        hipCheck(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)), "hipMalloc");
        n = count;
    }
    // grows to at least `count` elements (contents are not kept)
    void reserve(size_t count)
    {
        if (count > n) alloc(count);
    }
    void upload(const std::vector<T> &h, hipStream_t s)
    {
        alloc(h.size());
        if (!h.empty()) hipCheck(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
};

}  // namespace spgvbd
