// Split-precision radix-16 passes on gfx950's K = 32 matrix instructions, inside the single-plane
// Stockham chain (precision form, second generation).
//
// A radix-16 pass maps the 16 complex inputs x[r] of every butterfly to Y[q] = sum_r W[q][r] x[r],
// W[q][r] = e^{s 2 pi i q r / 16}. With the 32 real inputs of a butterfly interleaved (row 2r holds
// Re x[r], row 2r + 1 holds Im x[r]) one block of 16 butterflies (16 columns) is two GEMMs of K = 32:
//     Re Y = A_r . B,  A_r[q][2r] = Re W[q][r],  A_r[q][2r + 1] = -Im W[q][r]
//     Im Y = A_i . B,  A_i[q][2r] = Im W[q][r],  A_i[q][2r + 1] =  Re W[q][r]
// each one v_mfma_f32_16x16x32_{f16,bf16} per product of split terms. Lane layout of that
// instruction (checked on gfx950 with a one-wave probe): lane l holds A[l % 16][8 (l / 16) + t] and
// B[8 (l / 16) + t][l % 16] for t < 8, and D[4 (l / 16) + i][l % 16] for i < 4. So lane (g, j) =
// (l / 16, l % 16) holds the complex inputs r = 4g .. 4g + 3 of butterfly j of the block and gets its
// outputs q = 4g .. 4g + 3 back: every lane owns whole complex values, and loads the four pass
// twiddles W^{k r} it needs from a [k][16] table.
//
// Operands are split into FP16 / BF16 terms, x = x_hi + x_lo (+ x_mid for three-term BF16), the DFT
// matrix likewise (in FP64 on the host), and the products that matter are summed with FP32
// accumulation: hi.hi + hi.lo + lo.hi (two terms), plus hi.mid' + mid'.hi + mid.mid (three terms).
// FP16 has a narrow range, so each butterfly column is scaled by a power of two (max |x| into
// [2^13, 2^14)) before the split and unscaled after the GEMM, exactly.
//
// A wave owns 4 blocks (64 butterflies): block u of wave w covers the tile's butterflies
// (4 w + u) 16 + 0..15 (transform = index / T, butterfly = index mod T), so a launch needs whole
// waves (F T a multiple of 64) and T = L / 16.
#pragma once
#include <hip/hip_runtime.h>

#include "stockham_device.h"

namespace sip {

typedef _Float16 mf2_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mf2_b8 __attribute__((ext_vector_type(8)));
typedef float mf2_f4 __attribute__((ext_vector_type(4)));

enum Mf2Dtype { MF2_F16 = 1, MF2_BF16 = 2 };

template <int DT>
struct Mf2;
template <>
struct Mf2<MF2_F16> {
    typedef mf2_h8 vec;
    typedef _Float16 elem;
    static __device__ __forceinline__ mf2_f4 mma(vec a, vec b, mf2_f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mf2<MF2_BF16> {
    typedef mf2_b8 vec;
    typedef __bf16 elem;
    static __device__ __forceinline__ mf2_f4 mma(vec a, vec b, mf2_f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// This lane's A fragments: [term] of A_r and A_i, from the host table [term][part][lane][8].
template <int DT, int NT>
struct MfW2 {
    typename Mf2<DT>::vec ar[NT], ai[NT];
    __device__ __forceinline__ void load(const unsigned short *__restrict__ tab, int lane)
    {
#pragma unroll
        for (int tm = 0; tm < NT; tm++) {
            const uint4 pr = *reinterpret_cast<const uint4 *>(tab + ((tm * 2 + 0) * 64 + lane) * 8);
            const uint4 pi = *reinterpret_cast<const uint4 *>(tab + ((tm * 2 + 1) * 64 + lane) * 8);
            ar[tm] = __builtin_bit_cast(typename Mf2<DT>::vec, pr);
            ai[tm] = __builtin_bit_cast(typename Mf2<DT>::vec, pi);
        }
    }
};

// 8 reals (after `scale`) into NT terms, round to nearest even at every step.
template <int DT, int NT>
__device__ __forceinline__ void mf2_split(const float (&x)[8], float scale, typename Mf2<DT>::vec (&out)[NT])
{
    typedef typename Mf2<DT>::elem E;
    float r[8];
#pragma unroll
    for (int k = 0; k < 8; k++) r[k] = x[k] * scale;
#pragma unroll
    for (int tm = 0; tm < NT; tm++) {
#pragma unroll
        for (int k = 0; k < 8; k++) out[tm][k] = (E)r[k];
        if (tm + 1 < NT) {
#pragma unroll
            for (int k = 0; k < 8; k++) r[k] -= (float)out[tm][k];
        }
    }
}

// One block: x = this lane's inputs r = 4g .. 4g + 3 (twiddled), y = its outputs q = 4g .. 4g + 3.
template <int DT, int NT>
__device__ __forceinline__ void mf2_block(const MfW2<DT, NT> &W, const float2 (&x)[4], float2 (&y)[4])
{
    float scale = 1.f, unscale = 1.f;
    if constexpr (DT == MF2_F16) {
        // per-column power-of-two scale; the column lives in lanes j, j + 16, j + 32, j + 48 (the
        // maxima of |x| are then compared as bits: unsigned order is float order for non-negative
        // floats)
        const float m01 = fmaxf(fmaxf(fabsf(x[0].x), fabsf(x[0].y)), fabsf(x[1].x));
        const float m23 = fmaxf(fmaxf(fabsf(x[1].y), fabsf(x[2].x)), fabsf(x[2].y));
        unsigned mb = __float_as_uint(fmaxf(fmaxf(fabsf(x[3].x), fabsf(x[3].y)), fmaxf(m01, m23)));
        {
            const auto r = __builtin_amdgcn_permlane16_swap(mb, mb, false, false);
            mb = max((unsigned)r[0], (unsigned)r[1]);
        }
        {
            const auto r = __builtin_amdgcn_permlane32_swap(mb, mb, false, false);
            mb = max((unsigned)r[0], (unsigned)r[1]);
        }
        if (mb != 0u) {
            int e = (int)(mb >> 23) - 127;  // max |x| in [2^e, 2^(e+1))
            e = e < -100 ? -100 : (e > 100 ? 100 : e);
            scale = __uint_as_float((unsigned)(127 + 13 - e) << 23);
            unscale = __uint_as_float((unsigned)(127 - 13 + e) << 23);
        }
    }
    const float xf[8] = {x[0].x, x[0].y, x[1].x, x[1].y, x[2].x, x[2].y, x[3].x, x[3].y};
    typename Mf2<DT>::vec b[NT];
    mf2_split<DT, NT>(xf, scale, b);
    mf2_f4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
    if constexpr (NT == 3) {
        yr = Mf2<DT>::mma(W.ar[1], b[1], yr);
        yi = Mf2<DT>::mma(W.ai[1], b[1], yi);
        yr = Mf2<DT>::mma(W.ar[0], b[2], yr);
        yi = Mf2<DT>::mma(W.ai[0], b[2], yi);
        yr = Mf2<DT>::mma(W.ar[2], b[0], yr);
        yi = Mf2<DT>::mma(W.ai[2], b[0], yi);
    }
    yr = Mf2<DT>::mma(W.ar[1], b[0], yr);
    yi = Mf2<DT>::mma(W.ai[1], b[0], yi);
    yr = Mf2<DT>::mma(W.ar[0], b[1], yr);
    yi = Mf2<DT>::mma(W.ai[0], b[1], yi);
    yr = Mf2<DT>::mma(W.ar[0], b[0], yr);
    yi = Mf2<DT>::mma(W.ai[0], b[0], yi);
#pragma unroll
    for (int t = 0; t < 4; t++) y[t] = make_float2(yr[t] * unscale, yi[t] * unscale);
}

// The lane's twiddles W^{k r}, r = 4g + t (t < 4), from a [k][16] pass table (every r of a
// radix-16 pass stored): two 16-byte loads, no products.
__device__ __forceinline__ void mf2_twiddles(const float2 *__restrict__ tw, int k, int g, float2 (&w)[4])
{
    const float4 *p = reinterpret_cast<const float4 *>(tw + k * 16 + 4 * g);
    const float4 a = p[0], c = p[1];
    w[0] = make_float2(a.x, a.y);
    w[1] = make_float2(a.z, a.w);
    w[2] = make_float2(c.x, c.y);
    w[3] = make_float2(c.z, c.w);
}

// Rows per k of a pass table in a matrix chain: radix-16 passes keep all 16 W^{k r}.
template <int R>
constexpr int mf2_tw_rows()
{
    return R == 16 ? 16 : tw_rows<R>();
}

// Roles of a lane in the matrix layout: block u covers butterfly b[u] of tile transform fl[u].
// Rows (COLS = false): the 16 columns of a block are 16 consecutive butterflies, block u of wave w
// holding tile butterflies (4 w + u) 16 + 0..15 (transform = index / T); when T >= 64 a wave lies in
// one transform (FAST). Column tiles (COLS = true, F a multiple of 16): the 16 columns of a block
// are 16 adjacent transforms at one butterfly, block index 4 w + u = b (F / 16) + transform group,
// so that a wave's HBM accesses cover whole lines and its LDS accesses (transforms LP apart, LP
// odd) fall in different banks.
template <int T, bool COLS>
struct MfLane {
    static constexpr bool FAST = !COLS && T >= 64;
    int fl[4], b[4];
    int g, j;
    __device__ __forceinline__ MfLane(int tid, int F)
    {
        const int wave = tid >> 6, lane = tid & 63;
        g = lane >> 4;
        j = lane & 15;
        const int nfg = F >> 4;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if constexpr (COLS) {
                const int beta = wave * 4 + u;
                b[u] = beta / nfg;
                fl[u] = (beta - b[u] * nfg) * 16 + j;
            } else {
                const int gb = (wave * 4 + u) * 16 + j;
                fl[u] = gb / T;
                b[u] = gb - fl[u] * T;
            }
        }
    }
};

// Single-plane chain with radix-16 passes on the matrix cores (MF) and the others on the VALU.
// A radix-16 pass's register array v[0][4 u + t] holds input r = 4g + t of block u (matrix layout);
// every other pass uses the VALU layout of split_chain (thread t of transform lbase / LP, its
// butterflies t + q T). The exchanges move the outputs of one layout into the inputs of the next
// through the plane (real parts, then imaginary parts). io: load / store of transform fl (tile
// index) element j, and active(fl).
template <int L, int T, int S, int DT, int NT, int NS, int R, int... Rest, class Io, class Ml>
__device__ __forceinline__ void mf2_chain(float *pl, const float2 *__restrict__ tw, const MfW2<DT, NT> &W, int LP,
                                          int lbase, int t, bool active, const Ml &ml,
                                          float2 (&v)[(L / R + T - 1) / T][R], const Io &io)
{
    constexpr bool FAST = Ml::FAST;
    constexpr bool MF = R == 16;
    constexpr int NBF = L / R;
    constexpr int NB = (NBF + T - 1) / T;
    constexpr bool EXACT = (NBF % T) == 0;
    static_assert(!MF || (NB == 1 && NBF == T), "matrix passes need T = L / 16");
    // ---- this pass: twiddle + DFT in registers
    if constexpr (MF) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            float2 x[4], y[4];
#pragma unroll
            for (int tt = 0; tt < 4; tt++) x[tt] = v[0][4 * u + tt];
            if constexpr (NS > 1) {
                // W^0 = 1 is stored exactly, so lane group 0's first product is exact
                float2 w[4];
                mf2_twiddles(tw, ml.b[u] % NS, ml.g, w);
#pragma unroll
                for (int tt = 0; tt < 4; tt++) x[tt] = cmul(x[tt], w[tt]);
            }
            mf2_block<DT, NT>(W, x, y);
#pragma unroll
            for (int tt = 0; tt < 4; tt++) v[0][4 * u + tt] = y[tt];
        }
    } else {
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int b = t + q * T;
            if (active && (EXACT || b < NBF)) {
                if constexpr (NS > 1) {
                    float2 w[R];
                    tw_load<R, NS>(tw, b % NS, w);
                    tw_expand<R>(w);
#pragma unroll
                    for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], w[r]);
                }
                dft<R, S>(v[q]);
            }
        }
    }
    // ---- outputs: last pass to HBM, else through the plane into the next pass's layout
    if constexpr (sizeof...(Rest) == 0) {
        if constexpr (MF) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int b = ml.b[u], k = b % NS, d = (b / NS) * NS * 16 + k;
                if constexpr (FAST) {
                    // a wave lies inside one transform: the lane's own row
                    if (!active) continue;
#pragma unroll
                    for (int tt = 0; tt < 4; tt++) io.store_own(d + (4 * ml.g + tt) * NS, v[0][4 * u + tt]);
                } else {
                    if (!io.active(ml.fl[u])) continue;
#pragma unroll
                    for (int tt = 0; tt < 4; tt++) io.store(ml.fl[u], d + (4 * ml.g + tt) * NS, v[0][4 * u + tt]);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const int b = t + q * T;
                if (active && (EXACT || b < NBF)) {
                    const int k = b % NS, d = (b / NS) * NS * R + k;
#pragma unroll
                    for (int r = 0; r < R; r++) io.store_own(d + r * NS, v[q][r]);
                }
            }
        }
    } else {
        constexpr int R2 = first_of<Rest...>();
        constexpr bool MF2 = R2 == 16;
        constexpr int NBF2 = L / R2;
        constexpr int NB2 = (NBF2 + T - 1) / T;
        constexpr bool EXACT2 = (NBF2 % T) == 0;
        float2 n[NB2][R2];
#pragma unroll
        for (int plane = 0; plane < 2; plane++) {
            if (plane == 1 || NS > 1) __syncthreads();  // previous readers of the plane are done
            if constexpr (MF) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (!(FAST ? active : io.active(ml.fl[u]))) continue;
                    const int b = ml.b[u], k = b % NS, hi = b / NS, lb = FAST ? lbase : ml.fl[u] * LP;
#pragma unroll
                    for (int tt = 0; tt < 4; tt++)
                        pl[lb + lds_idx_write<NS, 16>(hi, k, 4 * ml.g + tt)] =
                            plane ? v[0][4 * u + tt].y : v[0][4 * u + tt].x;
                }
            } else {
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    const int b = t + q * T;
                    if (active && (EXACT || b < NBF)) {
                        const int k = b % NS, hi = b / NS;
#pragma unroll
                        for (int r = 0; r < R; r++)
                            pl[lbase + lds_idx_write<NS, R>(hi, k, r)] = plane ? v[q][r].y : v[q][r].x;
                    }
                }
            }
            __syncthreads();
            if constexpr (MF2) {
                // every lane reads (tiles hold whole waves; a block beyond the transforms of the launch
                // reads its own plane slots, whatever they hold, and its results are never stored)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int lb = FAST ? lbase : ml.fl[u] * LP;
#pragma unroll
                    for (int tt = 0; tt < 4; tt++) {
                        const float x = pl[lb + lds_idx_read<NBF2>(ml.b[u], 4 * ml.g + tt)];
                        if (plane)
                            n[0][4 * u + tt].y = x;
                        else
                            n[0][4 * u + tt].x = x;
                    }
                }
            } else {
#pragma unroll
                for (int q = 0; q < NB2; q++) {
                    const int b = t + q * T;
                    if (active && (EXACT2 || b < NBF2)) {
#pragma unroll
                        for (int r = 0; r < R2; r++) {
                            const float x = pl[lbase + lds_idx_read<NBF2>(b, r)];
                            if (plane)
                                n[q][r].y = x;
                            else
                                n[q][r].x = x;
                        }
                    }
                }
            }
        }
        mf2_chain<L, T, S, DT, NT, NS * R, Rest...>(pl, tw + NS * mf2_tw_rows<R>(), W, LP, lbase, t, active, ml, n, io);
    }
}

// Contiguous rows (or gathered rows) of a tile: element j of tile transform fl.
struct Mf2RowIo {
    const float2 *in;
    float2 *out;
    const StockhamArgs *a;
    int64_t tile;
    const float2 *own_in;
    float2 *own_out;
    __device__ __forceinline__ bool active(int fl) const { return tile * a->F + fl < a->nfft; }
    __device__ __forceinline__ int64_t in_base(int fl) const { return tile_base((uint32_t)(tile * a->F + fl), a->tin); }
    __device__ __forceinline__ int64_t out_base(int fl) const
    {
        return tile_base((uint32_t)(tile * a->F + fl), a->tout);
    }
    __device__ __forceinline__ void store(int fl, int j, float2 v) const
    {
        if (a->nt_out)
            gstore_nt(out + out_base(fl) + j, v);
        else
            gstore(out + out_base(fl) + j, v);
    }
    __device__ __forceinline__ void store_own(int j, float2 v) const
    {
        if (a->nt_out)
            gstore_nt(own_out + j, v);
        else
            gstore(own_out + j, v);
    }
};

// Single-plane rows kernel with the radix-16 passes on the matrix cores; the first pass reads HBM
// straight into its layout, the last writes HBM from its layout.
template <int L, int T, int S, int DT, int NT, int R0, int... Rest>
__device__ __forceinline__ void mf2_rows(const StockhamArgs &a, float *pl, int64_t tile)
{
    constexpr int NBF0 = L / R0;
    constexpr int NB0 = (NBF0 + T - 1) / T;
    constexpr bool EXACT0 = (NBF0 % T) == 0;
    const int F = a.F;
    const int tid = threadIdx.x;
    const int flp = tid / T;
    const int t = tid - flp * T;
    const MfLane<T, false> ml(tid, F);
    MfW2<DT, NT> W;
    W.load(a.mfma_w, tid & 63);
    Mf2RowIo io{static_cast<const float2 *>(a.in), static_cast<float2 *>(a.out), &a, tile, nullptr, nullptr};
    const bool active = io.active(flp);
    if (active) {
        io.own_in = io.in + io.in_base(flp);
        io.own_out = io.out + io.out_base(flp);
    }
    float2 v[NB0][R0];
    if constexpr (R0 == 16) {
        if constexpr (MfLane<T, false>::FAST) {
            // a wave lies inside one transform: the lane's own row (an inactive wave reads nothing)
            if (active && a.nt_in) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
#pragma unroll
                    for (int tt = 0; tt < 4; tt++)
                        v[0][4 * u + tt] = gload_nt(io.own_in + ml.b[u] + (4 * ml.g + tt) * NBF0);
                }
            } else if (active) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
#pragma unroll
                    for (int tt = 0; tt < 4; tt++) v[0][4 * u + tt] = gload(io.own_in + ml.b[u] + (4 * ml.g + tt) * NBF0);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                // a block beyond the launch's transforms re-reads the last transform (never stored)
                const int fl = io.active(ml.fl[u]) ? ml.fl[u] : (int)(a.nfft - 1 - tile * F);
                const float2 *g = io.in + io.in_base(fl) + ml.b[u];
#pragma unroll
                for (int tt = 0; tt < 4; tt++) v[0][4 * u + tt] = gload(g + (4 * ml.g + tt) * NBF0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload(io.own_in + b + r * NBF0);
            }
        }
    }
    (void)F;
    mf2_chain<L, T, S, DT, NT, 1, R0, Rest...>(pl, a.tw, W, a.LP, flp * a.LP, t, active, ml, v, io);
}

// Five waves per SIMD: the register budget (at most 102) that the heaviest instance (4096, three
// BF16 terms) meets without scratch; six would spill.
#ifndef SIP_MF2_WAVES
#define SIP_MF2_WAVES 5
#endif
template <int L, int T, int S, int DT, int NT, int... Rs>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(SIP_MF2_WAVES))) void stockham_mf2_rows_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    mf2_rows<L, T, S, DT, NT, Rs...>(a, lds, tile_index(a));
}

// Column tile (transforms adjacent, element j of transform fl at base(fl) + j es): element access
// for the matrix layout (any tile transform) and the VALU layout (the thread's own transform).
struct Mf2ColIo {
    const float2 *in;
    float2 *out;
    const StockhamArgs *a;
    int64_t tile;
    const float2 *own_in;
    float2 *own_out;
    __device__ __forceinline__ bool active(int fl) const { return tile * a->F + fl < a->nfft; }
    __device__ __forceinline__ const float2 *in_row(int fl) const
    {
        return in + tile_base((uint32_t)(tile * a->F + fl), a->tin);
    }
    __device__ __forceinline__ void store(int fl, int j, float2 v) const
    {
        gstore(out + tile_base((uint32_t)(tile * a->F + fl), a->tout) + (int64_t)j * a->tout.es, v);
    }
    __device__ __forceinline__ void store_own(int j, float2 v) const { gstore(own_out + (int64_t)j * a->tout.es, v); }
};

// Column tiles (no four-step twiddle) with the radix-16 passes on the matrix cores: the VALU passes
// walk the tile with lanes on adjacent transforms (thread tid: transform tid mod F, thread tid / F),
// the matrix passes with the column layout of MfLane; both first and last pass touch HBM directly.
template <int L, int T, int S, int DT, int NT, int R0, int... Rest>
__device__ __forceinline__ void mf2_cols(const StockhamArgs &a, float *pl, int64_t tile)
{
    constexpr int NBF0 = L / R0;
    constexpr int NB0 = (NBF0 + T - 1) / T;
    constexpr bool EXACT0 = (NBF0 % T) == 0;
    const int F = a.F;
    const int tid = threadIdx.x;
    const int t = (int)fdiv((uint32_t)tid, a.div_F);
    const int flp = tid - t * F;
    const MfLane<T, true> ml(tid, F);
    MfW2<DT, NT> W;
    W.load(a.mfma_w, tid & 63);
    Mf2ColIo io{static_cast<const float2 *>(a.in), static_cast<float2 *>(a.out), &a, tile, nullptr, nullptr};
    const bool active = io.active(flp);
    if (active) {
        io.own_in = io.in_row(flp);
        io.own_out = io.out + tile_base((uint32_t)(tile * F + flp), a.tout);
    }
    const int64_t esi = a.tin.es;
    float2 v[NB0][R0];
    if constexpr (R0 == 16) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            // a column beyond the launch's transforms re-reads the last transform (never stored)
            const int fl = io.active(ml.fl[u]) ? ml.fl[u] : (int)(a.nfft - 1 - tile * F);
            const float2 *g = io.in_row(fl) + (int64_t)ml.b[u] * esi;
#pragma unroll
            for (int tt = 0; tt < 4; tt++) v[0][4 * u + tt] = gload(g + (int64_t)((4 * ml.g + tt) * NBF0) * esi);
        }
    } else {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload(io.own_in + (int64_t)(b + r * NBF0) * esi);
            }
        }
    }
    mf2_chain<L, T, S, DT, NT, 1, R0, Rest...>(pl, a.tw, W, a.LP, flp * a.LP, t, active, ml, v, io);
}

// Column tiles with three BF16 terms: four waves per SIMD (128 VGPRs); at five they spill.
template <int L, int T, int S, int DT, int NT, int... Rs>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(NT == 3 ? 4 : SIP_MF2_WAVES))) void stockham_mf2_cols_kernel(
    const StockhamArgs a)
{
    extern __shared__ float lds[];
    mf2_cols<L, T, S, DT, NT, Rs...>(a, lds, tile_index(a));
}

}  // namespace sip
