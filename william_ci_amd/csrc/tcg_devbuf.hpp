// Owning device buffer of the host drivers (tcg_host.cpp, tcg_df.cpp, tcg_scf.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace tcg {

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // grows the buffer; the old one is freed only after the device has drained
    // (work queued earlier, without a synchronisation in between, may still read
    // it).  A grown buffer has a new address: callers re-read p after alloc.
    hipError_t alloc(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (p) {
            const hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) return e;
        }
        release();
        n = count;
        return hipMalloc((void **)&p, sizeof(T) * (count ? count : 1));
    }
    hipError_t upload(const std::vector<T> &v) {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess) return e;
        if (!v.empty()) e = hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
        return e;
    }
};

}  // namespace tcg
