// Fused 2D tile instances: X, Y in {32, 64, 128} (the reference's DD core range), with the radix
// sequences and threads per transform of the 1D leaves of those lengths.
#include "stockham2d.h"

namespace sip {

#define SIP_ARG(...) __VA_ARGS__
#define SIP_T2(X, TX, RXT, Y, TY, RYT)                                                               \
    {X, TX, Y, TY, Tile2d<X, TX, Y, TY, -1, RXT, RYT>::kThreads,                                     \
     (int)(2 * Tile2d<X, TX, Y, TY, -1, RXT, RYT>::kPlane * sizeof(float)),                          \
     (const void *)&stockham2d_kernel<X, TX, Y, TY, -1, RXT, RYT>,                                   \
     (const void *)&stockham2d_kernel<X, TX, Y, TY, 1, RXT, RYT>}

#define SIP_T2S(X, TX, RXT, Y, TY, RYT)                                                              \
    {X, TX, Y, TY, Tile2dSplit<X, TX, Y, TY, -1, RXT, RYT>::kThreads,                                \
     (int)(Tile2dSplit<X, TX, Y, TY, -1, RXT, RYT>::kPlane * sizeof(float)),                         \
     (const void *)&stockham2d_split_kernel<X, TX, Y, TY, -1, RXT, RYT>,                              \
     (const void *)&stockham2d_split_kernel<X, TX, Y, TY, 1, RXT, RYT>}

// single-plane 128 x 128 slice (70 KiB of LDS instead of 140 KiB)
extern const Stockham2dInstance g_stockham2d_split[1] = {
    SIP_T2S(128, 8, SIP_ARG(Radices<16, 8>), 128, 8, SIP_ARG(Radices<16, 8>)),
};

extern const Stockham2dInstance g_stockham2d[9] = {
    SIP_T2(32, 4, SIP_ARG(Radices<8, 4>), 32, 4, SIP_ARG(Radices<8, 4>)),
    SIP_T2(32, 4, SIP_ARG(Radices<8, 4>), 64, 8, SIP_ARG(Radices<8, 8>)),
    SIP_T2(32, 4, SIP_ARG(Radices<8, 4>), 128, 8, SIP_ARG(Radices<16, 8>)),
    SIP_T2(64, 8, SIP_ARG(Radices<8, 8>), 32, 4, SIP_ARG(Radices<8, 4>)),
    SIP_T2(64, 8, SIP_ARG(Radices<8, 8>), 64, 8, SIP_ARG(Radices<8, 8>)),
    SIP_T2(64, 8, SIP_ARG(Radices<8, 8>), 128, 8, SIP_ARG(Radices<16, 8>)),
    SIP_T2(128, 8, SIP_ARG(Radices<16, 8>), 32, 4, SIP_ARG(Radices<8, 4>)),
    SIP_T2(128, 8, SIP_ARG(Radices<16, 8>), 64, 8, SIP_ARG(Radices<8, 8>)),
    SIP_T2(128, 8, SIP_ARG(Radices<16, 8>), 128, 8, SIP_ARG(Radices<16, 8>)),
};

}  // namespace sip
