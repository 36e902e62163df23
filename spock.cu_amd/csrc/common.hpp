// Host-side pieces shared by the C ABI translation units: the context, device buffers and
// HIP error checks.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "handles.hpp"

struct spk_ctx : spk::Counted {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    double tickRateHz = 1e8;
};

namespace spk {

// Frees a context whose last reference was dropped: its stream, then the object.
inline void freeCtx(spk_ctx *c) {
    if (c->stream) {
        (void) hipStreamSynchronize(c->stream);
        (void) hipStreamDestroy(c->stream);
    }
    delete c;
}

inline void releaseCtx(spk_ctx *c) { release(c, freeCtx); }

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

// A cooperative launch the device refused; it queued nothing and can be retried with one
// workgroup (the C ABI returns SPK_ERR_COOPERATIVE_LAUNCH for it).
struct CoopRefused : HipError {
    using HipError::HipError;
};

inline int fail(spk_ctx *ctx, const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return dynamic_cast<const CoopRefused *>(&e) ? -2 : -1;
}

}  // namespace spk
