// dispatch.hpp -- runtime (dimension, storage precision) -> template instance.
#pragma once
#include "internal.hpp"

#define LQCD_DISPATCH_S(L, ...)                                                     \
    do {                                                                            \
        if ((L)->precision == 1) {                                                  \
            using S = float;                                                        \
            __VA_ARGS__;                                                            \
        } else {                                                                    \
            using S = double;                                                       \
            __VA_ARGS__;                                                            \
        }                                                                           \
    } while (0)

#define LQCD_DISPATCH_DS(L, ...)                                                    \
    do {                                                                            \
        switch ((L)->g.dim) {                                                       \
        case 2: {                                                                   \
            constexpr int D = 2;                                                    \
            LQCD_DISPATCH_S(L, __VA_ARGS__);                                        \
        } break;                                                                    \
        case 3: {                                                                   \
            constexpr int D = 3;                                                    \
            LQCD_DISPATCH_S(L, __VA_ARGS__);                                        \
        } break;                                                                    \
        case 4: {                                                                   \
            constexpr int D = 4;                                                    \
            LQCD_DISPATCH_S(L, __VA_ARGS__);                                        \
        } break;                                                                    \
        default:                                                                    \
            return hipErrorInvalidValue;                                            \
        }                                                                           \
    } while (0)

namespace lqcd {
inline unsigned grid_for(int64_t n, int block = kBlock)
{
    return (unsigned)((n + block - 1) / block);
}
} // namespace lqcd
