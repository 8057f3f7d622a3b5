// handles.hpp -- the opaque handle types behind the C ABI's pointers.
#pragma once

#include "common.hpp"

struct salsaa_ctx {
    salsaa::Context c;
};
struct salsaa_key {
    salsaa::Key k;
};

// Checks the context, holds its lock until the calling entry point returns
// (recursive: entry points that call other entry points keep it), and makes
// its device current.  Use once per function scope.
#define CTX_OK(ctx)                                                          \
    if (!(ctx)) {                                                            \
        ::salsaa::set_error("null context");                                 \
        return SALSAA_ERR_ARG;                                               \
    }                                                                        \
    std::lock_guard<std::recursive_mutex> salsaa_ctx_guard_((ctx)->c.lock); \
    SALSAA_HIP(hipSetDevice((ctx)->c.device))

// CTX_OK for entry points whose kernels exist at d = 64 only: a context of
// another degree gets a clear error instead of a wrong layout.
#define CTX_D64(ctx, what)                                                                     \
    CTX_OK(ctx);                                                                               \
    if ((ctx)->c.d != ::salsaa::kD) {                                                          \
        ::salsaa::set_error(std::string(what) +                                                \
                            ": the device kernels of this entry point support d = 64 only "    \
                            "(the context has d = " +                                          \
                            std::to_string((ctx)->c.d) + ")");                                 \
        return SALSAA_ERR_ARG;                                                                 \
    }
