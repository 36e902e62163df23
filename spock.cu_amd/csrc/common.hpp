// Host-side pieces shared by the C ABI translation units: the context, device buffers and
// HIP error checks.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

struct spk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    double tickRateHz = 1e8;
};

namespace spk {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        check(hipMalloc(&p, b), "hipMalloc");
        bytes = b;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

inline int fail(spk_ctx *ctx, const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return -1;
}

}  // namespace spk
