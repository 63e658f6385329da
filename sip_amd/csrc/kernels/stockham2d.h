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

// Element u of thread tid in HBM order over an X x Y slice (Y contiguous): i = tid + u NT, x = i / Y,
// y = i mod Y. When NT is a multiple of Y and divides X Y (every instance: powers of two), x = x0 +
// u DX with DX = NT / Y and x0 = tid / Y < DX, y = tid mod Y, so the padded LDS indices below are one
// per-thread base plus a compile-time offset per u (no per-u address registers live across the
// radix passes).
template <int X, int Y, int NT>
struct SliceMap {
    static constexpr int kPer = (X * Y + NT - 1) / NT;
    static constexpr bool kExact = (NT % Y) == 0 && ((X * Y) % NT) == 0;
    static constexpr int DX = kExact ? NT / Y : 1;
    static_assert(!kExact || DX % 16 == 0 || 16 % DX == 0, "DX must be a power of two");
    int tid, x0, y0;
    __device__ __forceinline__ explicit SliceMap(int t) : tid(t), x0(t / Y), y0(t - (t / Y) * Y) {}
    __device__ __forceinline__ bool ok(int u) const { return kExact || tid + u * NT < X * Y; }
    // [x][y] layout with padded pitch P
    template <int P>
    __device__ __forceinline__ int xy(int u) const
    {
        if constexpr (kExact) {
            return x0 * P + lds_pad(y0) + u * DX * P;
        } else {
            const int i = tid + u * NT, x = i / Y;
            return x * P + lds_pad(i - x * Y);
        }
    }
    // [y][x] layout with padded pitch Q: pad(x0 + u DX) = pad(x0) + u DX + (u DX) / 16 for x0 < DX
    template <int Q>
    __device__ __forceinline__ int yx(int u) const
    {
        if constexpr (kExact) {
            return y0 * Q + lds_pad(x0) + u * DX + ((u * DX) >> 4);
        } else {
            const int i = tid + u * NT, x = i / Y;
            return (i - x * Y) * Q + lds_pad(x);
        }
    }
};

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
        const SliceMap<X, Y, kThreads> m(tid);
        // load [x][y]: consecutive lanes read consecutive elements, all loads in flight first
        {
            float2 v[kPer];
#pragma unroll
            for (int u = 0; u < kPer; u++)
                if (m.ok(u)) v[u] = (src + u * kThreads)[tid];
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                if (m.ok(u)) {
                    const int j = m.template xy<kLpy>(u);
                    re[j] = v[u].x;
                    im[j] = v[u].y;
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
                if (m.ok(u)) {
                    const int j = m.template xy<kLpy>(u);
                    v[u] = make_float2(re[j], im[j]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                if (m.ok(u)) {
                    const int j = m.template yx<kLpx>(u);
                    re[j] = v[u].x;
                    im[j] = v[u].y;
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
            if (m.ok(u)) {
                const int j = m.template yx<kLpx>(u);
                (dst + u * kThreads)[tid] = make_float2(re[j], im[j]);
            }
        }
    }
};

template <int A, int... Rest>
constexpr int last_of()
{
    if constexpr (sizeof...(Rest) == 0)
        return A;
    else
        return last_of<Rest...>();
}

// The same slice with one LDS plane (half the LDS, so that a 128 x 128 slice leaves room for a
// second workgroup on the CU): every hand-over (HBM order -> row passes -> columns -> HBM order)
// moves the real parts through the plane, then the imaginary parts, and the radix passes of each
// axis are the single-plane chains of the 1D row kernel.
template <int X, int TX, int Y, int TY, int S, class RX, class RY>
struct Tile2dSplit;

template <int X, int TX, int Y, int TY, int S, int... RX, int... RY>
struct Tile2dSplit<X, TX, Y, TY, S, Radices<RX...>, Radices<RY...>> {
    static constexpr int kThreads = (X * TY > Y * TX) ? X * TY : Y * TX;
    static constexpr int kLpy = Y + Y / 16 + 1;  // padded row pitch, [x][y] layout (odd)
    static constexpr int kLpx = X + X / 16 + 1;  // padded column pitch, [y][x] layout (odd)
    static constexpr int kPlane = (X * kLpy > Y * kLpx) ? X * kLpy : Y * kLpx;
    static constexpr int kPer = (X * Y + kThreads - 1) / kThreads;  // elements per thread
    // first / last radix of each axis and the register shapes of those passes
    static constexpr int RY0 = first_of<RY...>(), RYL = last_of<RY...>();
    static constexpr int RX0 = first_of<RX...>(), RXL = last_of<RX...>();
    static constexpr int NBY0 = (Y / RY0 + TY - 1) / TY, NBYL = (Y / RYL + TY - 1) / TY;
    static constexpr int NBX0 = (X / RX0 + TX - 1) / TX, NBXL = (X / RXL + TX - 1) / TX;
    static constexpr int NSY = Y / RYL, NSX = X / RXL;  // NS of the last pass of each axis

    __device__ __forceinline__ static void run(const StockhamArgs &a, float *pl)
    {
        const int tid = threadIdx.x;
        const int64_t slice = blockIdx.x;
        const float2 *src = static_cast<const float2 *>(a.in) + slice * (int64_t)(X * Y);
        float2 *dst = static_cast<float2 *>(a.out) + slice * (int64_t)(X * Y);
        // row view: transform xr (length Y), thread tr; column view: transform yc (length X), tc
        const int xr = tid / TY, tr = tid - xr * TY;
        const bool ract = xr < X;
        const int yc = tid / TX, tc = tid - yc * TX;
        const bool cact = yc < Y;
        const int rbase = (ract ? xr : 0) * kLpy, cbase = (cact ? yc : 0) * kLpx;

        // HBM [x][y] -> first row pass: consecutive lanes read consecutive elements
        const SliceMap<X, Y, kThreads> m(tid);
        float2 rin[NBY0][RY0];
        {
            float2 v[kPer];
#pragma unroll
            for (int u = 0; u < kPer; u++)
                if (m.ok(u)) v[u] = (src + u * kThreads)[tid];
#pragma unroll
            for (int plane = 0; plane < 2; plane++) {
                if (plane) __syncthreads();
#pragma unroll
                for (int u = 0; u < kPer; u++)
                    if (m.ok(u)) pl[m.template xy<kLpy>(u)] = plane ? v[u].y : v[u].x;
                __syncthreads();
#pragma unroll
                for (int q = 0; q < NBY0; q++) {
                    const int b = tr + q * TY;
                    if (ract && b < Y / RY0) {
#pragma unroll
                        for (int r = 0; r < RY0; r++) {
                            const float x = pl[rbase + lds_idx_read<Y / RY0>(b, r)];
                            if (plane)
                                rin[q][r].y = x;
                            else
                                rin[q][r].x = x;
                        }
                    }
                }
            }
            __syncthreads();
        }
        // row passes
        float2 rout[NBYL][RYL];
        split_chain<Y, TY, S, 1, RY...>(pl, a.tw, rbase, tr, ract, rin, RegTail<RYL>{rout});
        // row outputs (x = xr, y = j) -> [y][x] -> first column pass
        float2 cin[NBX0][RX0];
#pragma unroll
        for (int plane = 0; plane < 2; plane++) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NBYL; q++) {
                const int b = tr + q * TY;
                if (ract && b < Y / RYL) {
                    const int k = b % NSY;
                    const int d = (b / NSY) * NSY * RYL + k;
#pragma unroll
                    for (int r = 0; r < RYL; r++) {
                        const int y = d + r * NSY;
                        pl[y * kLpx + lds_pad(xr)] = plane ? rout[q][r].y : rout[q][r].x;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NBX0; q++) {
                const int b = tc + q * TX;
                if (cact && b < X / RX0) {
#pragma unroll
                    for (int r = 0; r < RX0; r++) {
                        const float x = pl[cbase + lds_idx_read<X / RX0>(b, r)];
                        if (plane)
                            cin[q][r].y = x;
                        else
                            cin[q][r].x = x;
                    }
                }
            }
        }
        __syncthreads();
        // column passes
        float2 cout[NBXL][RXL];
        split_chain<X, TX, S, 1, RX...>(pl, a.tw2, cbase, tc, cact, cin, RegTail<RXL>{cout});
        // column outputs (y = yc, x = j) -> [x][y] -> HBM in order
        float2 v[kPer];
#pragma unroll
        for (int plane = 0; plane < 2; plane++) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NBXL; q++) {
                const int b = tc + q * TX;
                if (cact && b < X / RXL) {
                    const int k = b % NSX;
                    const int d = (b / NSX) * NSX * RXL + k;
#pragma unroll
                    for (int r = 0; r < RXL; r++) {
                        const int x = d + r * NSX;
                        pl[x * kLpy + lds_pad(yc)] = plane ? cout[q][r].y : cout[q][r].x;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                if (m.ok(u)) {
                    const float val = pl[m.template xy<kLpy>(u)];
                    if (plane)
                        v[u].y = val;
                    else
                        v[u].x = val;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; u++)
            if (m.ok(u)) (dst + u * kThreads)[tid] = v[u];
    }
};

template <int X, int TX, int Y, int TY, int S, class RX, class RY>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void stockham2d_split_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    Tile2dSplit<X, TX, Y, TY, S, RX, RY>::run(a, lds);
}

template <int X, int TX, int Y, int TY, int S, class RX, class RY>
__global__ __launch_bounds__(1024) void stockham2d_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    Tile2d<X, TX, Y, TY, S, RX, RY>::run(a, lds);
}

}  // namespace sip
