// Fused 2D tile kernel: a whole X x Y slice (Y contiguous) per workgroup, both axes transformed in
// LDS with one HBM read and one HBM write of the slice (the row pass over Y, a register transpose,
// the column pass over X). This is the job of the reference's DD core for 2D slices of 32..128
// per side (core/fft/fftcore/dd_core.cpp:26-141, P X Q products on the cube unit); here the small
// DFTs are the Stockham radix passes of the 1D kernels.
#pragma once
#include "stockham_device.h"

namespace sip {

template <int... R>
struct Radices {};

template <int X, int TX, int Y, int TY, int S, class RX, class RY>
struct Tile2d;

template <int X, int TX, int Y, int TY, int S, int... RX, int... RY>
struct Tile2d<X, TX, Y, TY, S, Radices<RX...>, Radices<RY...>> {
    static constexpr int kThreads = (X * TY > Y * TX) ? X * TY : Y * TX;
    static constexpr int kLpy = Y + Y / 16 + 1;  // padded row pitch, [x][y] layout (odd)
    static constexpr int kLpx = X + X / 16 + 1;  // padded column pitch, [y][x] layout (odd)
    static constexpr int kPlane = (X * kLpy > Y * kLpx) ? X * kLpy : Y * kLpx;
    static constexpr int kPer = (X * Y + kThreads - 1) / kThreads;  // elements per thread

    __device__ __forceinline__ static void run(const StockhamArgs &a, float *lds)
    {
        float *re = lds, *im = lds + kPlane;
        const int tid = threadIdx.x;
        const int64_t slice = blockIdx.x;
        const float2 *src = static_cast<const float2 *>(a.in) + slice * (int64_t)(X * Y);
        float2 *dst = static_cast<float2 *>(a.out) + slice * (int64_t)(X * Y);
        // load [x][y]: consecutive lanes read consecutive elements, all loads in flight first
        {
            float2 v[kPer];
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int i = tid + u * kThreads;
                if (i < X * Y) v[u] = src[i];
            }
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int i = tid + u * kThreads;
                if (i < X * Y) {
                    const int x = i / Y, y = i - x * Y;
                    re[x * kLpy + lds_pad(y)] = v[u].x;
                    im[x * kLpy + lds_pad(y)] = v[u].y;
                }
            }
        }
        __syncthreads();
        // rows: X transforms of length Y, TY threads each
        {
            const int x = tid / TY, t = tid - x * TY;
            const bool active = x < X;
            const RowIo none{nullptr, nullptr};
            passes_static<Y, TY, S, false, false, 1, RY...>(re, im, a.tw, (active ? x : 0) * kLpy, t, active, none);
        }
        // register transpose [x][y] -> [y][x]
        {
            float2 v[kPer];
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int i = tid + u * kThreads;
                if (i < X * Y) {
                    const int x = i / Y, y = i - x * Y;
                    v[u] = make_float2(re[x * kLpy + lds_pad(y)], im[x * kLpy + lds_pad(y)]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int i = tid + u * kThreads;
                if (i < X * Y) {
                    const int x = i / Y, y = i - x * Y;
                    re[y * kLpx + lds_pad(x)] = v[u].x;
                    im[y * kLpx + lds_pad(x)] = v[u].y;
                }
            }
            __syncthreads();
        }
        // columns: Y transforms of length X, TX threads each
        {
            const int y = tid / TX, t = tid - y * TX;
            const bool active = y < Y;
            const RowIo none{nullptr, nullptr};
            passes_static<X, TX, S, false, false, 1, RX...>(re, im, a.tw2, (active ? y : 0) * kLpx, t, active, none);
        }
        // store (x, y) from [y][x]: consecutive lanes write consecutive y
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int i = tid + u * kThreads;
            if (i < X * Y) {
                const int x = i / Y, y = i - x * Y;
                dst[i] = make_float2(re[y * kLpx + lds_pad(x)], im[y * kLpx + lds_pad(x)]);
            }
        }
    }
};

template <int X, int TX, int Y, int TY, int S, class RX, class RY>
__global__ __launch_bounds__(1024) void stockham2d_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    Tile2d<X, TX, Y, TY, S, RX, RY>::run(a, lds);
}

}  // namespace sip
