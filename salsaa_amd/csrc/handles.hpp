// handles.hpp -- the opaque handle types behind the C ABI's pointers.
#pragma once

#include "common.hpp"

struct salsaa_ctx {
    salsaa::Context c;
};
struct salsaa_key {
    salsaa::Key k;
};

#define CTX_OK(ctx)                                                  \
    do {                                                             \
        if (!(ctx)) {                                                \
            ::salsaa::set_error("null context");                     \
            return SALSAA_ERR_ARG;                                   \
        }                                                            \
        SALSAA_HIP(hipSetDevice((ctx)->c.device));                   \
    } while (0)
