// Split-precision radix-16 Stockham stage on the matrix cores (precision form).
//
// A radix-16 stage maps 16 complex inputs of every butterfly to 16 outputs with the DFT matrix
// W[q][r] = e^{s 2 pi i q r / 16}. In real form that is one GEMM per block of 16 butterflies:
//     [Y_re; Y_im] (32 x 16) = [[W_re, -W_im], [W_im, W_re]] (32 x 32) . [X_re; X_im] (32 x 16),
// run as 2 x 2 v_mfma_f32_16x16x16 tiles (A = DFT matrix, B = butterfly columns, FP32 accumulate).
// FP16 / BF16 operands lose precision, so both operands are split: x = x_hi + x_lo (+ x_mid for
// three-term BF16), w likewise (split in FP64 on the host), and the products that matter are
// summed: hi*hi + hi*lo + lo*hi (two terms) or hi*hi + hi*mid + mid*hi + hi*lo + lo*hi + mid*mid
// (three terms), which brings the stage back to about FP32 accuracy. FP16 has a narrow range: each
// butterfly column is scaled by a power of two (max |x| -> [2^13, 2^14)) before the split and
// unscaled after the GEMM, exactly.
//
// Lane layout of v_mfma_f32_16x16x16 (f16 and bf16_1k), verified on gfx950: lane l holds
// A[l % 16][4 (l / 16) + t], B[4 (l / 16) + t][l % 16] and D[4 (l / 16) + t][l % 16], t = 0..3.
// With T = L / 16 threads per transform each wave owns exactly 4 blocks of 16 butterflies.
#pragma once
#include <hip/hip_runtime.h>

#include "stockham_device.h"

namespace sip {

typedef _Float16 mf_half4 __attribute__((ext_vector_type(4)));
typedef short mf_short4 __attribute__((ext_vector_type(4)));
typedef float mf_float4 __attribute__((ext_vector_type(4)));

enum MfmaDtype { MF_F16 = 1, MF_BF16 = 2 };

template <int DT>
struct MfOp;
template <>
struct MfOp<MF_F16> {
    typedef mf_half4 vec;
    static __device__ __forceinline__ unsigned short to_bits(float x)
    {
        const _Float16 h = (_Float16)x;
        return __builtin_bit_cast(unsigned short, h);
    }
    static __device__ __forceinline__ float from_bits(unsigned short b)
    {
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ mf_float4 mma(vec a, vec b, mf_float4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ vec pack(const unsigned short *v)
    {
        vec r;
#pragma unroll
        for (int t = 0; t < 4; t++) r[t] = __builtin_bit_cast(_Float16, v[t]);
        return r;
    }
};
template <>
struct MfOp<MF_BF16> {
    typedef mf_short4 vec;
    static __device__ __forceinline__ unsigned short to_bits(float x)
    {
        // round to nearest even (the gfx950 conversion instruction)
        const __bf16 h = (__bf16)x;
        return __builtin_bit_cast(unsigned short, h);
    }
    static __device__ __forceinline__ float from_bits(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
    static __device__ __forceinline__ mf_float4 mma(vec a, vec b, mf_float4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ vec pack(const unsigned short *v)
    {
        vec r;
#pragma unroll
        for (int t = 0; t < 4; t++) r[t] = (short)v[t];
        return r;
    }
};

// DFT-matrix operands of one lane: [term][rb][kb] (A fragments), loaded once per kernel from the
// host-built table laid out [term][rb][kb][lane][4] (16-bit values).
template <int DT, int NT>
struct MfW {
    typename MfOp<DT>::vec a[NT][2][2];
    __device__ __forceinline__ void load(const unsigned short *__restrict__ tab, int lane)
    {
#pragma unroll
        for (int tm = 0; tm < NT; tm++)
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    const unsigned short *p = tab + ((((tm * 2 + rb) * 2 + kb) * 64 + lane) * 4);
                    unsigned short v[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) v[t] = p[t];
                    a[tm][rb][kb] = MfOp<DT>::pack(v);
                }
    }
};

// Splits the 4 values of a lane's column slice into NT terms (optionally after scaling).
template <int DT, int NT>
__device__ __forceinline__ void mf_split(const float (&x)[4], float scale, typename MfOp<DT>::vec (&out)[NT])
{
    unsigned short b[NT][4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float r = x[t] * scale;
#pragma unroll
        for (int tm = 0; tm < NT; tm++) {
            b[tm][t] = MfOp<DT>::to_bits(r);
            r -= MfOp<DT>::from_bits(b[tm][t]);
        }
    }
#pragma unroll
    for (int tm = 0; tm < NT; tm++) out[tm] = MfOp<DT>::pack(b[tm]);
}

// Twiddle rows of a pass in an MFMA chain: radix-16 passes keep every W^{k r}, r = 0..15, k-major
// ([k][16]), so that each lane loads the four it multiplies (r = 4 gq .. 4 gq + 3) with two 16-byte
// loads; the VALU passes keep the plain layout.
template <int R>
constexpr int mf_tw_rows()
{
    return R == 16 ? 16 : tw_rows<R>();
}

// One radix-16 pass on MFMA for a tile of F transforms of length L in LDS (T = L/16 threads per
// transform, F * T a multiple of 64). NS = product of the earlier radices. GIN: the pass reads its
// inputs from HBM (src + in_base[f]); GOUT: it writes its outputs to HBM (dst + out_base[f]).
template <int L, int T, int S, int NS, int DT, int NT, bool GIN, bool GOUT>
__device__ __forceinline__ void pass_mfma16(float *re, float *im, const float2 *__restrict__ tw, const MfW<DT, NT> &W,
                                            int LP, int nf, const float2 *__restrict__ src, float2 *__restrict__ dst,
                                            const int64_t *in_base, const int64_t *out_base)
{
    constexpr int NBF = L / 16;  // butterflies per transform (== T)
    static_assert(NBF == T, "MFMA stage expects T = L / 16");
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int j = lane & 15, gq = lane >> 4;
    float xr[4][4], xi[4][4];
    int fls[4], bs[4];
    // gather: block u of this wave covers global butterflies (4 wave + u) * 16 + 0..15. Every lane
    // gathers (fl < F always, so LDS reads stay inside the tile; HBM reads of transforms beyond nf
    // re-read the last valid one); only the stores are guarded.
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int g = (wave * 4 + u) * 16 + j;
        const int fl = g / NBF;
        const int b = g - fl * NBF;
        fls[u] = fl;
        bs[u] = b;
        // this lane's twiddles W^{k r}, r = 4 gq + t, straight from the [k][16] pass table
        float2 lane_tw[4];
        if constexpr (NS > 1) {
            const float4 *tp = reinterpret_cast<const float4 *>(tw + (b % NS) * 16 + 4 * gq);
            const float4 p0 = tp[0], p1 = tp[1];
            lane_tw[0] = make_float2(p0.x, p0.y);
            lane_tw[1] = make_float2(p0.z, p0.w);
            lane_tw[2] = make_float2(p1.x, p1.y);
            lane_tw[3] = make_float2(p1.z, p1.w);
        }
        const float2 *gsrc = nullptr;
        if constexpr (GIN) gsrc = src + in_base[fl < nf ? fl : nf - 1] + b;
        const int lbase = fl * LP;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int r = 4 * gq + t;
            float vr, vi;
            if constexpr (GIN) {
                const float2 v = gload(gsrc + r * NBF);
                vr = v.x;
                vi = v.y;
            } else {
                const int jl = lds_idx_read<NBF>(b, r);
                vr = re[lbase + jl];
                vi = im[lbase + jl];
            }
            if constexpr (NS > 1) {
                const float2 p = cmul(make_float2(vr, vi), lane_tw[t]);  // W^0 = 1 exactly for r = 0
                vr = p.x;
                vi = p.y;
            }
            xr[u][t] = vr;
            xi[u][t] = vi;
        }
    }
    if constexpr (!GIN) __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
        float scale = 1.f, unscale = 1.f;
        if constexpr (DT == MF_F16) {
            // per-column power-of-two scale: the column lives in lanes j, j+16, j+32, j+48
            // max |x| as bits (unsigned order = float order for non-negative floats)
            unsigned mb = 0;
#pragma unroll
            for (int t = 0; t < 4; t++)
                mb = max(mb, max(__float_as_uint(xr[u][t]) & 0x7FFFFFFFu, __float_as_uint(xi[u][t]) & 0x7FFFFFFFu));
            // v_permlane16_swap with both operands = mb leaves rows 2i and 2i+1 of the two results
            // holding each other's values, v_permlane32_swap the two half-waves': two maxima give
            // every lane the maximum over lanes j, j+16, j+32, j+48
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
        typename MfOp<DT>::vec br[NT], bi[NT];
        mf_split<DT, NT>(xr[u], scale, br);
        mf_split<DT, NT>(xi[u], scale, bi);
        mf_float4 acc[2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            acc[rb] = mf_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                const auto *bb = kb == 0 ? br : bi;
                // hi*hi, hi*lo(.), lo(.)*hi and, for three terms, hi*lo, lo*hi, mid*mid
                acc[rb] = MfOp<DT>::mma(W.a[0][rb][kb], bb[0], acc[rb]);
                acc[rb] = MfOp<DT>::mma(W.a[0][rb][kb], bb[1], acc[rb]);
                acc[rb] = MfOp<DT>::mma(W.a[1][rb][kb], bb[0], acc[rb]);
                if constexpr (NT == 3) {
                    acc[rb] = MfOp<DT>::mma(W.a[0][rb][kb], bb[2], acc[rb]);
                    acc[rb] = MfOp<DT>::mma(W.a[2][rb][kb], bb[0], acc[rb]);
                    acc[rb] = MfOp<DT>::mma(W.a[1][rb][kb], bb[1], acc[rb]);
                }
            }
        }
        const int fl = fls[u];
        if (fl < nf) {
            const int b = bs[u];
            const int k = b % NS;
            const int hi = b / NS;
            const int lbase = fl * LP;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int q = 4 * gq + t;
                if constexpr (GOUT) {
                    gstore(dst + out_base[fl] + hi * NS * 16 + k + q * NS,
                           make_float2(acc[0][t] * unscale, acc[1][t] * unscale));
                } else {
                    const int jl = lds_idx_write<NS, 16>(hi, k, q);
                    re[lbase + jl] = acc[0][t] * unscale;
                    im[lbase + jl] = acc[1][t] * unscale;
                }
            }
        }
    }
    if constexpr (!GOUT) __syncthreads();
}

// Radix chain with radix-16 passes on MFMA and the others on the VALU (pass_static). DIN / DOUT:
// the first pass reads HBM / the last pass writes HBM directly.
struct MixedIo {
    const float2 *src;           // tile input (MFMA passes, per-block transform bases)
    float2 *dst;
    const int64_t *in_base, *out_base;
    const float2 *gin;           // this thread's transform (VALU passes)
    float2 *gout;
};

template <int L, int T, int S, int DT, int NT, bool DIN, bool DOUT, int NS, int R, int... Rest>
__device__ __forceinline__ void passes_mixed(float *re, float *im, const float2 *__restrict__ tw,
                                             const MfW<DT, NT> &W, int LP, int nf, int lbase, int t, bool active,
                                             const MixedIo &io)
{
    constexpr bool first = NS == 1, last = sizeof...(Rest) == 0;
    if constexpr (R == 16)
        pass_mfma16<L, T, S, NS, DT, NT, DIN && first, DOUT && last>(re, im, tw, W, LP, nf, io.src, io.dst,
                                                                     io.in_base, io.out_base);
    else
        pass_static<L, T, S, NS, R, DIN && first, DOUT && last>(re, im, tw, lbase, t, active, RowIo{io.gin, io.gout});
    if constexpr (!last)
        passes_mixed<L, T, S, DT, NT, DIN, DOUT, NS * R, Rest...>(re, im, tw + NS * mf_tw_rows<R>(), W, LP, nf, lbase, t,
                                                                 active, io);
}

template <int L, int T, int S, int DT, int NT, int... Rs>
struct MfmaRunner {
    __device__ __forceinline__ void operator()(const StockhamArgs &a, float *re, float *im, int lbase, int t,
                                               bool active, const float2 *gin, float2 *gout, int nf,
                                               const int64_t *in_base, const int64_t *out_base, int64_t) const
    {
        MfW<DT, NT> W;
        W.load(a.mfma_w, (int)(threadIdx.x & 63));
        const MixedIo io{static_cast<const float2 *>(a.in), static_cast<float2 *>(a.out), in_base, out_base,
                         gin, gout};
        if (a.direct_in && a.direct_out)
            passes_mixed<L, T, S, DT, NT, true, true, 1, Rs...>(re, im, a.tw, W, a.LP, nf, lbase, t, active, io);
        else if (a.direct_in)
            passes_mixed<L, T, S, DT, NT, true, false, 1, Rs...>(re, im, a.tw, W, a.LP, nf, lbase, t, active, io);
        else if (a.direct_out)
            passes_mixed<L, T, S, DT, NT, false, true, 1, Rs...>(re, im, a.tw, W, a.LP, nf, lbase, t, active, io);
        else
            passes_mixed<L, T, S, DT, NT, false, false, 1, Rs...>(re, im, a.tw, W, a.LP, nf, lbase, t, active, io);
    }
};

// Tile kernel with the radix-16 stages on the matrix cores.
template <int L, int T, int S, int DT, int NT, int... Rs>
__global__ __launch_bounds__(1024) void stockham_mfma_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    stockham_body<L, T, S>(a, lds, MfmaRunner<L, T, S, DT, NT, Rs...>{});
}

}  // namespace sip
