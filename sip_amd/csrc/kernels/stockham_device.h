// LDS-staged Stockham FFT kernels for gfx950 (the plain method).
//
// A workgroup owns F transforms of length L (a "tile"). The tile is read from HBM with the
// unit-stride axis walked by consecutive lanes (rows for contiguous batches, transforms for
// column passes), kept in LDS as separate re / im float planes padded by one word every 16
// (so both row-wise and column-wise lane patterns avoid bank conflicts), pushed through the
// radix passes, and written back the same way. Each pass is the Stockham autosort step:
// butterfly b reads elements b + r*L/R, multiplies by W_{ns R}^{(b mod ns) r}, runs an R-point DFT
// in registers and writes elements (b/ns)*ns*R + (b mod ns) + q*ns. The radix sequence, the
// thread count per transform and therefore every LDS index are compile-time constants of the
// kernel instance; only the tile addressing (strides, batch) is runtime.
// Twiddles come from FP64-generated tables rounded once to FP32.
//
// The reference computes the same recursion as staged DFT-matrix GEMMs on the Ascend cube unit
// (FFT-B four-step, ops/fft/c2c/fft_b/fft_b/kernel/fft_b.cce:2290-2576; FFT-Mix multi-step,
// ops/fft/common/include/kernel/fft_all_common.cceh:593-616); here the small DFTs run as VALU
// butterflies and the re-layouts between stages stay in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include "butterflies.h"
#include "stockham_args.h"

namespace sip {

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &d)
{
    return (__umulhi(n, d.m) + n) >> d.s;
}

__device__ __forceinline__ int lds_pad(int j) { return j + (j >> 4); }

// Streaming HBM accesses of the direct passes (read once / written once): nontemporal loads and
// stores are chosen per launch (StockhamArgs::nt_in / nt_out).
typedef float sip_vf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gload(const float2 *p) { return *p; }
__device__ __forceinline__ float2 gload_nt(const float2 *p)
{
    const sip_vf2 v = __builtin_nontemporal_load(reinterpret_cast<const sip_vf2 *>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gstore(float2 *p, float2 v) { *p = v; }
__device__ __forceinline__ void gstore_nt(float2 *p, float2 v)
{
    sip_vf2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<sip_vf2 *>(p));
}

template <int A, int... Rest>
constexpr int first_of()
{
    return A;
}

// Padded LDS index of element base + r * STRIDE for a compile-time STRIDE. When the padding of
// the r-dependent part is a compile-time constant the result is pad(base) + const(r), so a
// butterfly's R accesses share one address and differ only in the instruction's offset field.
template <int STRIDE>
__device__ __forceinline__ int lds_idx_read(int base, int r)
{
    if constexpr (STRIDE % 16 == 0)
        return lds_pad(base) + r * (STRIDE + STRIDE / 16);
    else
        return lds_pad(base + r * STRIDE);
}

// Padded index of the Stockham output element hi*NS*R + k + r*NS (k < NS).
template <int NS, int R>
__device__ __forceinline__ int lds_idx_write(int hi, int k, int r)
{
    if constexpr (NS % 16 == 0) {
        return lds_pad(hi * NS * R + k) + r * (NS + NS / 16);
    } else if constexpr (16 % NS == 0 && (NS * R) % 16 == 0) {
        // k + r*NS never crosses a multiple of 16 inside one r step because NS divides 16
        return hi * (NS * R + NS * R / 16) + k + r * NS + (r * NS) / 16;
    } else {
        return lds_pad(hi * NS * R + k + r * NS);
    }
}

// Twiddle rows of a radix-R pass: only W^{k r} for the powers of two r = 1, 2, 4, .. < R are stored
// (the other r are products of those: w[r] = w[highest power of two in r] * w[rest], at most four
// multiplies deep for r < 32), k-major and padded to an even count, so that a pass loads its
// twiddles with one to three 16-byte loads: [k][2] (radix 2, 3, 4), [k][4] (5 .. 16), [k][6] (20).
// Every pass has an even number of rows, so every pass table starts 16-byte aligned.
template <int R>
constexpr int tw_nlog()
{
    int n = 0;
    while ((1 << n) < R) n++;
    return n;
}

template <int R>
constexpr int tw_rows()
{
    return (tw_nlog<R>() + 1) & ~1;
}

template <int R, int NS>
__device__ __forceinline__ void tw_load(const float2 *__restrict__ tw, int k, float2 (&w)[R])
{
    constexpr int NL = tw_nlog<R>(), NR = tw_rows<R>();
    const float4 *p = reinterpret_cast<const float4 *>(tw + k * NR);
#pragma unroll
    for (int i = 0; i < NR / 2; i++) {
        const float4 a = p[i];
        if (2 * i < NL) w[1 << (2 * i)] = make_float2(a.x, a.y);
        if (2 * i + 1 < NL) w[1 << (2 * i + 1)] = make_float2(a.z, a.w);
    }
}

// Completes w[1 .. R-1] from the loaded powers of two.
template <int R>
__device__ __forceinline__ void tw_expand(float2 (&w)[R])
{
#pragma unroll
    for (int r = 3; r < R; r++) {
        if ((r & (r - 1)) == 0) continue;
        int hb = 1;
        while (hb * 2 <= r) hb *= 2;
        w[r] = cmul(w[hb], w[r - hb]);
    }
}

// HBM endpoints of the direct passes: element j of this thread's transform.
struct RowIo {  // contiguous transform
    static constexpr bool kPairTail = false;
    static constexpr bool kRegTail = false;
    const float2 *__restrict__ in;
    float2 *__restrict__ out;
    bool nt = false;  // nontemporal stores (an output that no later launch of the plan reads)
    __device__ __forceinline__ float2 load(int j) const { return gload(in + j); }
    __device__ __forceinline__ void store(int j, float2 v) const
    {
        if (nt)
            gstore_nt(out + j, v);
        else
            gstore(out + j, v);
    }
    // the R outputs d + r NS of one butterfly
    template <int R, int NS>
    __device__ __forceinline__ void store_bf(int d, const float2 (&v)[R]) const
    {
        if (nt) {
#pragma unroll
            for (int r = 0; r < R; r++) gstore_nt(out + d + r * NS, v[r]);
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) gstore(out + d + r * NS, v[r]);
        }
    }
};

// One radix-R pass over a transform of length L held in LDS, T threads per transform,
// NS = product of the radices of the earlier passes, TW = this pass's twiddle slots [k][tw_rows<R>].
// GIN: the pass reads its inputs straight from HBM (first pass: element b + r*L/R, through
// io.load) instead of LDS. GOUT: the pass writes its outputs straight to HBM (last pass: the
// autosort output index is b + r*L/R as well, through io.store).
template <int L, int T, int S, int NS, int R, bool GIN, bool GOUT, class Io>
__device__ __forceinline__ void pass_static(float *re, float *im, const float2 *__restrict__ tw, int lbase, int t,
                                            bool active, const Io &io)
{
    constexpr int NBF = L / R;                 // butterflies per transform
    constexpr int NB = (NBF + T - 1) / T;      // butterflies per thread
    constexpr bool EXACT = (NBF % T) == 0;
    float2 v[NB][R];
    float2 w[NB][R];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = t + q * T;
        if (active && (EXACT || b < NBF)) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                if constexpr (GIN) {
                    v[q][r] = io.load(b + r * NBF);
                } else {
                    const int j = lds_idx_read<NBF>(b, r);
                    v[q][r] = make_float2(re[lbase + j], im[lbase + j]);
                }
            }
            if constexpr (NS > 1) tw_load<R, NS>(tw, b % NS, w[q]);
        }
    }
    if constexpr (!GIN && !GOUT) __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = t + q * T;
        if (active && (EXACT || b < NBF)) {
            if constexpr (NS > 1) {
                tw_expand<R>(w[q]);
#pragma unroll
                for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], w[q][r]);
            }
            dft<R, S>(v[q]);
            const int k = b % NS;
            const int hi = b / NS;
            const int d = hi * NS * R + k;
            if constexpr (GOUT) {
                io.template store_bf<R, NS>(d, v[q]);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int j = lds_idx_write<NS, R>(hi, k, r);
                    re[lbase + j] = v[q][r].x;
                    im[lbase + j] = v[q][r].y;
                }
            }
        }
    }
    if constexpr (!GOUT) __syncthreads();
}

template <int L, int T, int S, bool DIN, bool DOUT, int NS, int R, int... Rest, class Io>
__device__ __forceinline__ void passes_static(float *re, float *im, const float2 *__restrict__ tw, int lbase, int t,
                                              bool active, const Io &io)
{
    constexpr bool first = NS == 1;
    constexpr bool last = sizeof...(Rest) == 0;
    pass_static<L, T, S, NS, R, DIN && first, DOUT && last>(re, im, tw, lbase, t, active, io);
    if constexpr (!last)
        passes_static<L, T, S, DIN, DOUT, NS * R, Rest...>(re, im, tw + NS * tw_rows<R>(), lbase, t, active, io);
}

// f = f0 + n0 * (f1 + n1 * f2)
__device__ __forceinline__ void tile_split(uint32_t f, const TileDesc &d, uint32_t &f0, uint32_t &f1, uint32_t &f2)
{
    const uint32_t g = fdiv(f, d.div_n0);
    f0 = f - g * (uint32_t)d.n0;
    f2 = fdiv(g, d.div_n1);
    f1 = g - f2 * (uint32_t)d.n1;
}

// Element offset of transform f's element 0 (strided or gathered through the row table).
__device__ __forceinline__ int64_t tile_base(uint32_t f, const TileDesc &d, uint32_t *f0_out = nullptr)
{
    uint32_t f0, f1, f2;
    tile_split(f, d, f0, f1, f2);
    if (f0_out) *f0_out = f0;
    int64_t o0 = (int64_t)f0 * d.s0, o1 = (int64_t)f1 * d.s1, o2 = (int64_t)f2 * d.s2;
    if (d.tab) {
        if (d.tab_dim == 0) o0 = d.tab[f0];
        else if (d.tab_dim == 1) o1 = d.tab[f1];
        else o2 = d.tab[f2];
    }
    return o0 + o1 + o2;
}

// W^m with W = e^{s 2 pi i / Ntot}, from the two-level table (m < Ntot).
__device__ __forceinline__ float2 tw4_root(const StockhamArgs &a, int64_t m)
{
    return cmul(a.tw4_lo[m & a.tw4_mask], a.tw4_hi[m >> a.tw4_shift]);
}

// Real-transform pair arithmetic (half-length complex transform Z of the even / odd samples,
// w^k = e^{s 2 pi i k / 2L}):
//   R2C bin  X[k] = Z[k] (1 - i w^k)/2 + conj Z[L-k] (1 + i w^k)/2,
//   C2R input P[k] = (X[k] + conj X[L-k]) + i w^k (X[k] - conj X[L-k]).
// Bin / input L - k uses w^{L-k} = -conj w^k with the roles of k and L - k exchanged.
__device__ __forceinline__ float2 r2c_bin(float2 z, float2 zc, float2 w)
{
    const float2 iw = mul_si<1>(w);
    const float2 A = make_float2(0.5f * (1.f - iw.x), -0.5f * iw.y);
    const float2 B = make_float2(0.5f * (1.f + iw.x), 0.5f * iw.y);
    return cadd(cmul(z, A), cmul(cconj(zc), B));
}
__device__ __forceinline__ float2 c2r_input(float2 x, float2 xc, float2 w)
{
    const float2 c = cconj(xc);
    return cadd(cadd(x, c), mul_si<1>(cmul(w, csub(x, c))));
}

// R2C end of a single-plane chain: the last pass's outputs Z (natural order) pass through the
// plane to the post-process units (unit kk > 0: bins kk and L - kk from Z[kk], Z[L-kk]; unit 0:
// bins 0 and L from Z[0]; for even L thread 0 also writes the self-paired bin L/2).
struct R2cTail {
    static constexpr bool kPairTail = true;
    static constexpr bool kRegTail = false;
    float2 *__restrict__ out;  // this transform's L + 1 bins
    const float2 *__restrict__ rtw;
};

template <int L, int T, int NS, int R>
__device__ __forceinline__ void r2c_tail(float *pl, int lbase, int t, bool active,
                                         const float2 (&v)[(L / R + T - 1) / T][R], const R2cTail &io)
{
    constexpr int NBF = L / R;
    constexpr int NB = (NBF + T - 1) / T;
    constexpr bool EXACT = (NBF % T) == 0;
    constexpr int H = L / 2;
    constexpr int NUN = (L + 1) / 2;
    constexpr int NU = (NUN + T - 1) / T;
    constexpr bool EVEN = (L % 2) == 0;
    float2 wk[NU], zk[NU], zm[NU], zh = make_float2(0.f, 0.f), wh = make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int kk = t + u * T;
        if (active && kk < NUN) wk[u] = io.rtw[kk];
    }
    if (EVEN && active && t == 0) wh = io.rtw[H];
#pragma unroll
    for (int plane = 0; plane < 2; plane++) {
        __syncthreads();  // earlier readers of the plane are done
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int b = t + q * T;
            if (active && (EXACT || b < NBF)) {
                const int k = b % NS;
                const int hi = b / NS;
#pragma unroll
                for (int r = 0; r < R; r++) pl[lbase + lds_idx_write<NS, R>(hi, k, r)] = plane ? v[q][r].y : v[q][r].x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int kk = t + u * T;
            if (active && kk < NUN) {
                const int km = kk == 0 ? 0 : L - kk;
                const float x0 = pl[lbase + lds_pad(kk)], x1 = pl[lbase + lds_pad(km)];
                if (plane) {
                    zk[u].y = x0;
                    zm[u].y = x1;
                } else {
                    zk[u].x = x0;
                    zm[u].x = x1;
                }
            }
        }
        if (EVEN && active && t == 0) {
            const float x = pl[lbase + lds_pad(H)];
            if (plane)
                zh.y = x;
            else
                zh.x = x;
        }
    }
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int kk = t + u * T;
        if (active && kk < NUN) {
            gstore(io.out + kk, r2c_bin(zk[u], zm[u], wk[u]));
            gstore(io.out + (L - kk), r2c_bin(zm[u], zk[u], make_float2(-wk[u].x, wk[u].y)));
        }
    }
    if (EVEN && active && t == 0) gstore(io.out + H, r2c_bin(zh, zh, wh));
}

// End of a single-plane chain that keeps the last pass's outputs in registers: o[q][r] holds
// element d + r NS of the transform (d = (b / NS) NS R + b mod NS, b = t + q T), for the caller to
// move on (the fused 2D tile's transposes).
template <int R>
struct RegTail {
    static constexpr bool kPairTail = false;
    static constexpr bool kRegTail = true;
    float2 (*o)[R];
};

// Thread roles of a single-plane chain. NoRemap: every pass keeps the thread's (transform, thread
// in transform). LastRemap: the exchange before the last pass hands the thread another transform
// and thread index (plane offset lbase, thread t), so the first and the last pass may walk the
// tile in different lane orders (rows in, columns out).
struct NoRemap {
    static constexpr bool kRemap = false;
    int lbase, t;
    bool active;
};
struct LastRemap {
    static constexpr bool kRemap = true;
    int lbase, t;
    bool active;
};

// C2C transforms with one LDS plane per transform (half the LDS of the two-plane layout, so
// twice as many workgroups fit on a CU): contiguous rows, or column tiles whose lanes walk adjacent
// columns. The first pass reads HBM, the last writes HBM through io.store, and every exchange in
// between moves the real parts through the plane, then the imaginary parts.
template <int L, int T, int S, int NS, int R, int... Rest, class Io, class Rm = NoRemap>
__device__ __forceinline__ void split_chain(float *pl, const float2 *__restrict__ tw, int lbase, int t, bool active,
                                            float2 (&v)[(L / R + T - 1) / T][R], const Io &io,
                                            const Rm &rm = Rm{0, 0, false})
{
    constexpr int NBF = L / R;
    constexpr int NB = (NBF + T - 1) / T;
    constexpr bool EXACT = (NBF % T) == 0;
    // this pass: twiddle + DFT in registers
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
    if constexpr (sizeof...(Rest) == 0 && Io::kPairTail) {
        r2c_tail<L, T, NS, R>(pl, lbase, t, active, v, io);
    } else if constexpr (sizeof...(Rest) == 0 && Io::kRegTail) {
#pragma unroll
        for (int q = 0; q < NB; q++)
#pragma unroll
            for (int r = 0; r < R; r++) io.o[q][r] = v[q][r];
    } else if constexpr (sizeof...(Rest) == 0) {
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int b = t + q * T;
            if (active && (EXACT || b < NBF)) {
                const int k = b % NS;
                const int d = (b / NS) * NS * R + k;
                io.template store_bf<R, NS>(d, v[q]);
            }
        }
    } else {
        constexpr int R2 = first_of<Rest...>();
        constexpr int NBF2 = L / R2;
        constexpr int NB2 = (NBF2 + T - 1) / T;
        constexpr bool EXACT2 = (NBF2 % T) == 0;
        // roles in the next pass (a different transform / thread before the last pass of a remap)
        constexpr bool kNow = Rm::kRemap && sizeof...(Rest) == 1;
        const int lbase2 = kNow ? rm.lbase : lbase, t2 = kNow ? rm.t : t;
        const bool active2 = kNow ? rm.active : active;
        float2 n[NB2][R2];
#pragma unroll
        for (int plane = 0; plane < 2; plane++) {
            if (plane == 1 || NS > 1) __syncthreads();  // previous readers of the plane are done
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const int b = t + q * T;
                if (active && (EXACT || b < NBF)) {
                    const int k = b % NS;
                    const int hi = b / NS;
#pragma unroll
                    for (int r = 0; r < R; r++)
                        pl[lbase + lds_idx_write<NS, R>(hi, k, r)] = plane ? v[q][r].y : v[q][r].x;
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NB2; q++) {
                const int b = t2 + q * T;
                if (active2 && (EXACT2 || b < NBF2)) {
#pragma unroll
                    for (int r = 0; r < R2; r++) {
                        const float x = pl[lbase2 + lds_idx_read<NBF2>(b, r)];
                        if (plane)
                            n[q][r].y = x;
                        else
                            n[q][r].x = x;
                    }
                }
            }
        }
        split_chain<L, T, S, NS * R, Rest...>(pl, tw + NS * tw_rows<R>(), lbase2, t2, active2, n, io, rm);
    }
}

template <int L, int T, int S, bool R2C, int R0, int... Rest>
__device__ __forceinline__ void split_rows(const StockhamArgs &a, float *pl, int64_t tile)
{
    constexpr int NBF0 = L / R0;
    constexpr int NB0 = (NBF0 + T - 1) / T;
    constexpr bool EXACT0 = (NBF0 % T) == 0;
    const int F = a.F;
    const int tid = threadIdx.x;
    const int flp = tid / T;
    const int t = tid - flp * T;
    const int64_t f = tile * F + flp;
    const bool active = f < a.nfft;
    int64_t ib = 0, ob = 0;
    if (active) {
        ib = tile_base((uint32_t)f, a.tin);
        ob = tile_base((uint32_t)f, a.tout);
    }
    const float2 *g = static_cast<const float2 *>(a.in) + ib;
    float2 v[NB0][R0];
    if (a.nt_in) {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload_nt(g + b + r * NBF0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload(g + b + r * NBF0);
            }
        }
    }
    if constexpr (R2C)
        split_chain<L, T, S, 1, R0, Rest...>(pl, a.tw, flp * a.LP, t, active, v,
                                             R2cTail{static_cast<float2 *>(a.out) + ob, a.rtw});
    else
        split_chain<L, T, S, 1, R0, Rest...>(pl, a.tw, flp * a.LP, t, active, v,
                                             RowIo{nullptr, static_cast<float2 *>(a.out) + ob, a.nt_out != 0});
}

// C2R rows with one LDS plane: the half spectrum X[0..L] is read from HBM in pair units (unit kk > 0
// reads X[kk], X[L-kk] and forms P[kk], P[L-kk]; unit 0 forms P[0] with the imaginary parts of X[0]
// and X[L] dropped; for even L thread 0 also forms the self-paired P[L/2]), exchanged through the
// plane into the first pass's layout, and the chain writes the L complex (2L real) outputs to HBM.
template <int L, int T, int S, int R0, int... Rest>
__device__ __forceinline__ void split_rows_c2r(const StockhamArgs &a, float *pl, int64_t tile)
{
    constexpr int NBF0 = L / R0;
    constexpr int NB0 = (NBF0 + T - 1) / T;
    constexpr bool EXACT0 = (NBF0 % T) == 0;
    constexpr int H = L / 2;
    constexpr int NUN = (L + 1) / 2;
    constexpr int NU = (NUN + T - 1) / T;
    constexpr bool EVEN = (L % 2) == 0;
    const int F = a.F;
    const int tid = threadIdx.x;
    const int flp = tid / T;
    const int t = tid - flp * T;
    const int64_t f = tile * F + flp;
    const bool active = f < a.nfft;
    int64_t ib = 0, ob = 0;
    if (active) {
        ib = tile_base((uint32_t)f, a.tin);
        ob = tile_base((uint32_t)f, a.tout);
    }
    const int lbase = flp * a.LP;
    const float2 *g = static_cast<const float2 *>(a.in) + ib;
    float2 p1[NU], p2[NU], ph = make_float2(0.f, 0.f);
    {
        float2 xk[NU], xm[NU], wk[NU], xh = make_float2(0.f, 0.f), wh = make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int kk = t + u * T;
            if (active && kk < NUN) {
                xk[u] = gload(g + kk);
                xm[u] = gload(g + (L - kk));
                wk[u] = a.rtw[kk];
            }
        }
        if (EVEN && active && t == 0) {
            xh = gload(g + H);
            wh = a.rtw[H];
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int kk = t + u * T;
            if (active && kk < NUN) {
                float2 x0 = xk[u], x1 = xm[u];
                if (kk == 0) {
                    x0.y = 0.f;
                    x1.y = 0.f;
                }
                p1[u] = c2r_input(x0, x1, wk[u]);
                p2[u] = c2r_input(x1, x0, make_float2(-wk[u].x, wk[u].y));
            }
        }
        if (EVEN && active && t == 0) ph = c2r_input(xh, xh, wh);
    }
    float2 v[NB0][R0];
#pragma unroll
    for (int plane = 0; plane < 2; plane++) {
        if (plane == 1) __syncthreads();  // plane 0's readers are done
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int kk = t + u * T;
            if (active && kk < NUN) {
                pl[lbase + lds_pad(kk)] = plane ? p1[u].y : p1[u].x;
                if (kk != 0) pl[lbase + lds_pad(L - kk)] = plane ? p2[u].y : p2[u].x;
            }
        }
        if (EVEN && active && t == 0) pl[lbase + lds_pad(H)] = plane ? ph.y : ph.x;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) {
                    const float x = pl[lbase + lds_idx_read<NBF0>(b, r)];
                    if (plane)
                        v[q][r].y = x;
                    else
                        v[q][r].x = x;
                }
            }
        }
    }
    __syncthreads();  // the first pass writes the plane again
    split_chain<L, T, S, 1, R0, Rest...>(pl, a.tw, lbase, t, active, v,
                                         RowIo{nullptr, static_cast<float2 *>(a.out) + ob});
}

// Tile of this workgroup; with xcd_remap the round-robin dispatch of consecutive workgroups to the
// 8 XCDs is undone so that each XCD streams through one contiguous range of tiles.
__device__ __forceinline__ int64_t tile_index(const StockhamArgs &a)
{
    int64_t tile = blockIdx.x;
    if (a.xcd_remap) {
        const int64_t ntiles = gridDim.x, q = ntiles / 8, rem = ntiles % 8;
        const int64_t x = tile % 8, sl = tile / 8;
        tile = x * q + (x < rem ? x : rem) + sl;
    }
    return tile;
}

// Contiguous C2C rows, single LDS plane (separate kernel so that its register budget is its own).
template <int L, int T, int S, int... Rs>
__global__ __launch_bounds__(1024) void stockham_rows_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    split_rows<L, T, S, false, Rs...>(a, lds, tile_index(a));
}

// Real-transform rows, single LDS plane: MODE 1 = R2C (half-length C2C + pair post-process),
// MODE 2 = C2R (pair pre-process + half-length C2C).
template <int L, int T, int S, int MODE, int... Rs>
__global__ __launch_bounds__(1024) void stockham_real_rows_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    if constexpr (MODE == 1)
        split_rows<L, T, S, true, Rs...>(a, lds, tile_index(a));
    else
        split_rows_c2r<L, T, S, Rs...>(a, lds, tile_index(a));
}

// Strided transform of a column tile: element j at in[j * es_in] / out[j * es_out], with the
// four-step twiddle W^{col * j} applied on store when the stage carries one. TW: 0 = the stage has
// no four-step twiddle, 1 = it has one, 2 = decided at run time (tw4_ntot). The single-plane
// column kernels are instantiated for 0 and 1 so that the twiddle code does not raise the register
// budget of the plain column passes (cols 1024: 61 VGPRs without, 76 with).
template <int TW>
struct ColIoT {
    static constexpr bool kPairTail = false;
    static constexpr bool kRegTail = false;
    const float2 *__restrict__ in;
    float2 *__restrict__ out;
    int64_t es_in, es_out;
    int64_t col;
    const StockhamArgs *a;
    __device__ __forceinline__ float2 load(int j) const { return gload(in + (int64_t)j * es_in); }
    __device__ __forceinline__ void store(int j, float2 v) const
    {
        if (TW == 1 || (TW == 2 && a->tw4_ntot != 0)) v = cmul(v, tw4_root(*a, col * (int64_t)j));
        gstore(out + (int64_t)j * es_out, v);
    }
    // The R outputs d + r NS of one butterfly. With a four-step twiddle, W^{col (d + r NS)} =
    // W^{col d} (W^{col NS})^r: W^{col d} and the powers-of-two r of W^{col NS} come from the
    // two-level table (every argument is below col * L < ntot), the other r are multiplied
    // together as for the pass twiddles, so a radix-16 butterfly loads 10 table entries instead
    // of 32.
    template <int R, int NS>
    __device__ __forceinline__ void store_bf(int d, const float2 (&v)[R]) const
    {
        if (TW == 0 || (TW == 2 && a->tw4_ntot == 0)) {
            if (a->nt_out) {
#pragma unroll
                for (int r = 0; r < R; r++) gstore_nt(out + (int64_t)(d + r * NS) * es_out, v[r]);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) gstore(out + (int64_t)(d + r * NS) * es_out, v[r]);
            }
            return;
        }
        const int64_t step = col * (int64_t)NS;
        float2 w[R];
#pragma unroll
        for (int i = 0; (1 << i) < R; i++) w[1 << i] = tw4_root(*a, step << i);
        const float2 base = tw4_root(*a, col * (int64_t)d);
        tw_expand<R>(w);
        gstore(out + (int64_t)d * es_out, cmul(v[0], base));
#pragma unroll
        for (int r = 1; r < R; r++) gstore(out + (int64_t)(d + r * NS) * es_out, cmul(v[r], cmul(base, w[r])));
    }
};
using ColIo = ColIoT<2>;

// Column tile with one LDS plane per transform: lane tid walks transform tid mod F (adjacent
// columns, so a wave's HBM accesses cover whole lines) and butterflies tid / F; the first pass
// reads and the last pass writes HBM through ColIo (with the four-step twiddle on store).
template <int L, int T, int S, bool TW4, int R0, int... Rest>
__device__ __forceinline__ void split_cols(const StockhamArgs &a, float *pl, int64_t tile)
{
    constexpr int NBF0 = L / R0;
    constexpr int NB0 = (NBF0 + T - 1) / T;
    constexpr bool EXACT0 = (NBF0 % T) == 0;
    const int F = a.F;
    const int tid = threadIdx.x;
    const int t = (int)fdiv((uint32_t)tid, a.div_F);
    const int flp = tid - t * F;
    const int64_t f = tile * F + flp;
    const bool active = f < a.nfft;
    int64_t ib = 0, ob = 0, col = 0;
    if (active) {
        uint32_t f0;
        ib = tile_base((uint32_t)f, a.tin, &f0);
        col = fdiv(f0, a.div_col);
        ob = tile_base((uint32_t)f, a.tout);
    }
    const ColIoT<TW4 ? 1 : 0> io{static_cast<const float2 *>(a.in) + ib, static_cast<float2 *>(a.out) + ob,
                                 a.tin.es, a.tout.es, col, &a};
    float2 v[NB0][R0];
    if (a.nt_in) {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload_nt(io.in + (int64_t)(b + r * NBF0) * io.es_in);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = t + q * T;
            if (active && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = io.load(b + r * NBF0);
            }
        }
    }
    split_chain<L, T, S, 1, R0, Rest...>(pl, a.tw, flp * a.LP, t, active, v, io);
}

// Column tiles, single LDS plane (separate kernel so that its register budget is its own); TW4:
// the stage applies the four-step twiddle on store.
template <int L, int T, int S, bool TW4, int... Rs>
__global__ __launch_bounds__(1024) void stockham_cols_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    split_cols<L, T, S, TW4, Rs...>(a, lds, tile_index(a));
}

// Rows in, columns out (the second stage of a four-step split: transform k1 is a contiguous row of
// the intermediate and its element k2 goes to k1 + L1 k2, so adjacent transforms are adjacent on
// output). The first pass reads each row straight from HBM with the transform's T threads on
// consecutive elements (thread tid: transform tid / T); the exchange before the last pass regroups
// the threads so that lanes walk adjacent transforms (thread tid: transform tid mod F), and the last
// pass writes straight to HBM, a whole run of F adjacent outputs per element. Needs two or more
// passes (the regrouping happens in an exchange).
template <int L, int T, int S, int R0, int... Rest>
__device__ __forceinline__ void split_rows_cols(const StockhamArgs &a, float *pl, int64_t tile)
{
    static_assert(sizeof...(Rest) >= 1, "rows-in / columns-out needs two or more passes");
    constexpr int NBF0 = L / R0;
    constexpr int NB0 = (NBF0 + T - 1) / T;
    constexpr bool EXACT0 = (NBF0 % T) == 0;
    const int F = a.F;
    const int tid = threadIdx.x;
    const int fr = tid / T, tr = tid - fr * T;  // row roles
    const int tc = (int)fdiv((uint32_t)tid, a.div_F);
    const int fc = tid - tc * F;                // column roles
    const int64_t f1 = tile * F + fr, f2 = tile * F + fc;
    const bool act1 = f1 < a.nfft, act2 = f2 < a.nfft;
    int64_t ib = 0, ob = 0;
    if (act1) ib = tile_base((uint32_t)f1, a.tin);
    if (act2) ob = tile_base((uint32_t)f2, a.tout);
    const float2 *g = static_cast<const float2 *>(a.in) + ib;
    float2 v[NB0][R0];
    if (a.nt_in) {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = tr + q * T;
            if (act1 && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload_nt(g + b + r * NBF0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < NB0; q++) {
            const int b = tr + q * T;
            if (act1 && (EXACT0 || b < NBF0)) {
#pragma unroll
                for (int r = 0; r < R0; r++) v[q][r] = gload(g + b + r * NBF0);
            }
        }
    }
    const ColIoT<0> io{nullptr, static_cast<float2 *>(a.out) + ob, 0, a.tout.es, 0, &a};
    split_chain<L, T, S, 1, R0, Rest...>(pl, a.tw, fr * a.LP, tr, act1, v, io, LastRemap{fc * a.LP, tc, act2});
}

template <int L, int T, int S, int... Rs>
__global__ __launch_bounds__(1024) void stockham_rc_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    split_rows_cols<L, T, S, Rs...>(a, lds, tile_index(a));
}

// Pass runners: the radix chain the tile goes through between load and store.
template <int L, int T, int S, int... Rs>
struct StaticRunner {
    __device__ __forceinline__ void operator()(const StockhamArgs &a, float *re, float *im, int lbase, int t,
                                               bool active, const float2 *gin, float2 *gout, int nf,
                                               const int64_t *, const int64_t *, int64_t col) const
    {
        const RowIo io{gin, gout};
        if (a.col_direct) {
            const ColIo cio{gin, gout, a.tin.es, a.tout.es, col, &a};
            passes_static<L, T, S, true, true, 1, Rs...>(re, im, a.tw, lbase, t, active, cio);
        } else if (a.direct_in && a.direct_out)
            passes_static<L, T, S, true, true, 1, Rs...>(re, im, a.tw, lbase, t, active, io);
        else if (a.direct_in)
            passes_static<L, T, S, true, false, 1, Rs...>(re, im, a.tw, lbase, t, active, io);
        else if (a.direct_out)
            passes_static<L, T, S, false, true, 1, Rs...>(re, im, a.tw, lbase, t, active, io);
        else
            passes_static<L, T, S, false, false, 1, Rs...>(re, im, a.tw, lbase, t, active, io);
    }
};

// Generic tile kernel body: per-transform bases, tile load (with C2R pre-processing), the runner's
// radix passes, tile store (with R2C post-processing / four-step twiddle).
template <int L, int T, int S, class Runner>
__device__ __forceinline__ void stockham_body(const StockhamArgs &a, float *lds, const Runner &run)
{
    const int nthreads = blockDim.x;
    const int tid = threadIdx.x;
    const int F = a.F;
    const int LP = a.LP;
    float *re = lds;
    float *im = lds + F * LP;
    int64_t *in_base = reinterpret_cast<int64_t *>(lds + 2 * F * LP);
    int64_t *out_base = in_base + F;
    int64_t *col = out_base + F;

    const int64_t tile = tile_index(a);

    const int64_t fstart = tile * F;
    const int nf = (int)((a.nfft - fstart) < F ? (a.nfft - fstart) : F);

    // Per-transform base offsets (and the four-step column index).
    for (int fl = tid; fl < nf; fl += nthreads) {
        const uint32_t f = (uint32_t)(fstart + fl);
        uint32_t f0;
        in_base[fl] = tile_base(f, a.tin, &f0);
        col[fl] = fdiv(f0, a.div_col);
        out_base[fl] = tile_base(f, a.tout);
    }
    __syncthreads();

    // ---------------------------------------------------------------- load tile
    const int in_len = a.in_len;
    if (a.direct_in || a.col_direct) {
        // the first radix pass reads HBM directly
    } else if (a.in_sep) {
        // split planes (C2C): real and imaginary parts from two float arrays
        const float *sre = static_cast<const float *>(a.in), *sim = a.in_im;
        const int total = F * in_len;
        for (int i = tid; i < total; i += nthreads) {
            int fl, j;
            if (a.tin.unit_f) {
                j = (int)fdiv((uint32_t)i, a.div_F);
                fl = i - j * F;
            } else {
                fl = (int)fdiv((uint32_t)i, a.div_in_len);
                j = i - fl * in_len;
            }
            if (fl < nf) {
                const int64_t o = in_base[fl] + (int64_t)j * a.tin.es;
                re[fl * LP + lds_pad(j)] = sre[o];
                im[fl * LP + lds_pad(j)] = sim[o];
            }
        }
    } else if (a.io_mode == IO_R2C_FULL) {
        const float *src = static_cast<const float *>(a.in);
        const int total = F * in_len;
        for (int i = tid; i < total; i += nthreads) {
            int fl, j;
            if (a.tin.unit_f) {
                j = (int)fdiv((uint32_t)i, a.div_F);
                fl = i - j * F;
            } else {
                fl = (int)fdiv((uint32_t)i, a.div_in_len);
                j = i - fl * in_len;
            }
            if (fl < nf) {
                re[fl * LP + lds_pad(j)] = src[in_base[fl] + (int64_t)j * a.tin.es];
                im[fl * LP + lds_pad(j)] = 0.f;
            }
        }
    } else if (a.io_mode == IO_C2R_EVEN) {
        // Half spectrum X[0..L] straight from HBM into the pre-processed L-point input
        // P[k] = (X[k] + conj X[L-k]) + i w^k (X[k] - conj X[L-k]) (imag of X[0] and X[L] dropped).
        // Unit kk > 0 of a transform reads X[kk] and X[L-kk] and writes P[kk] and P[L-kk]
        // (w^{L-k} = -conj w^k); unit 0 writes P[0]. For even L the self-paired P[L/2] is one more
        // step per transform. The block is F * T threads and a transform has (L+1)/2 units, so a
        // thread has at most NU; all loads come first.
        constexpr int H = L / 2;
        constexpr int NUN = (L + 1) / 2;
        constexpr bool EVEN = (L % 2) == 0;
        constexpr int NU = (NUN + T - 1) / T;
        const float2 *src = static_cast<const float2 *>(a.in);
        const int64_t es = a.tin.es;
        float2 xk[NU], xm[NU], wk[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int i = tid + u * nthreads;
            const int fl = i / NUN;
            const int kk = i - fl * NUN;
            if (i < F * NUN && fl < nf) {
                const float2 *x = src + in_base[fl];
                xk[u] = x[(int64_t)kk * es];
                xm[u] = x[(int64_t)(L - kk) * es];
                wk[u] = a.rtw[kk];
            }
        }
        if (EVEN && tid < nf) {
            const float2 xh = src[in_base[tid] + (int64_t)H * es];
            const float2 p = c2r_input(xh, xh, a.rtw[H]);
            re[tid * LP + lds_pad(H)] = p.x;
            im[tid * LP + lds_pad(H)] = p.y;
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int i = tid + u * nthreads;
            const int fl = i / NUN;
            const int kk = i - fl * NUN;
            if (i >= F * NUN || fl >= nf) continue;
            float *pr = re + fl * LP, *pi = im + fl * LP;
            float2 a0 = xk[u], a1 = xm[u];
            if (kk == 0) {
                a0.y = 0.f;
                a1.y = 0.f;
            }
            {
                const float2 p = c2r_input(a0, a1, wk[u]);
                pr[lds_pad(kk)] = p.x;
                pi[lds_pad(kk)] = p.y;
            }
            if (kk != 0) {
                const float2 p = c2r_input(a1, a0, make_float2(-wk[u].x, wk[u].y));
                pr[lds_pad(L - kk)] = p.x;
                pi[lds_pad(L - kk)] = p.y;
            }
        }
    } else if (a.tin.vec2) {
        // the block is F * T threads: every thread moves at most NV float4 (all loads in flight
        // before the first LDS store); in_len is at most L + 1 (C2R half spectrum)
        constexpr int NV = ((L + 1) / 2 + T - 1) / T;
        const float4 *src = static_cast<const float4 *>(a.in);
        const int half = in_len >> 1;
        const int total = F * half;
        float4 v[NV];
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int i = tid + u * nthreads;
            const int fl = (int)fdiv((uint32_t)i, a.div_in_vec);
            if (i < total && fl < nf) v[u] = src[(in_base[fl] >> 1) + (i - fl * half)];
        }
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int i = tid + u * nthreads;
            const int fl = (int)fdiv((uint32_t)i, a.div_in_vec);
            if (i < total && fl < nf) {
                const int j = 2 * (i - fl * half);
                re[fl * LP + lds_pad(j)] = v[u].x;
                im[fl * LP + lds_pad(j)] = v[u].y;
                re[fl * LP + lds_pad(j + 1)] = v[u].z;
                im[fl * LP + lds_pad(j + 1)] = v[u].w;
            }
        }
    } else if (a.tin.unit_f && a.col_fixed) {
        // column tile: each thread keeps one transform (fl = tid mod F) and loads its elements
        // j0, j0 + T, ... (the block is F * T threads), all loads in flight before the LDS stores
        constexpr int NU = (L + T - 1) / T;
        const int j0 = (int)fdiv((uint32_t)tid, a.div_F);
        const int fl = tid - j0 * F;
        if (fl < nf) {
            const float2 *p = static_cast<const float2 *>(a.in) + in_base[fl];
            const int64_t es = a.tin.es;
            float *pr = re + fl * LP, *pi = im + fl * LP;
            float2 v[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const int j = j0 + u * T;
                if (j < in_len) v[u] = p[(int64_t)j * es];
            }
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const int j = j0 + u * T;
                if (j < in_len) {
                    pr[lds_pad(j)] = v[u].x;
                    pi[lds_pad(j)] = v[u].y;
                }
            }
        }
    } else {
        // generic complex tile (e.g. the L+1 bins of a C2R row): in_len <= L + 1 and the block is
        // F * T threads, so a thread moves at most NG elements; all loads are issued first
        constexpr int NG = (L + 1 + T - 1) / T;
        const float2 *src = static_cast<const float2 *>(a.in);
        const int total = F * in_len;
        float2 v[NG];
        int dst[NG];
#pragma unroll
        for (int u = 0; u < NG; u++) {
            const int i = tid + u * nthreads;
            dst[u] = -1;
            if (i < total) {
                int fl, j;
                if (a.tin.unit_f) {
                    j = (int)fdiv((uint32_t)i, a.div_F);
                    fl = i - j * F;
                } else {
                    fl = (int)fdiv((uint32_t)i, a.div_in_len);
                    j = i - fl * in_len;
                }
                if (fl < nf) {
                    v[u] = src[in_base[fl] + (int64_t)j * a.tin.es];
                    dst[u] = fl * LP + lds_pad(j);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NG; u++) {
            if (dst[u] >= 0) {
                re[dst[u]] = v[u].x;
                im[dst[u]] = v[u].y;
            }
        }
    }
    __syncthreads();

    // ---------------------------------------------------------------- pre-process (C2R, odd)
    if (a.io_mode == IO_C2R_FULL) {
        // Hermitian extension of bins 0..L/2 to L points; imag of bin 0 (and L/2) dropped.
        constexpr int H = L / 2;
        for (int i = tid; i < F * L; i += nthreads) {
            const int fl = i / L;
            const int j = i - fl * L;
            if (fl >= nf) continue;
            float *pr = re + fl * LP, *pi = im + fl * LP;
            if (j > H) {
                pr[lds_pad(j)] = pr[lds_pad(L - j)];
                pi[lds_pad(j)] = -pi[lds_pad(L - j)];
            }
        }
        __syncthreads();
        for (int fl = tid; fl < nf; fl += nthreads) {
            im[fl * LP] = 0.f;
            if ((L & 1) == 0) im[fl * LP + lds_pad(H)] = 0.f;
        }
        __syncthreads();
    }

    // ---------------------------------------------------------------- radix passes
    {
        // rows: T consecutive threads per transform; column tiles with direct HBM passes:
        // consecutive lanes walk the F transforms (adjacent columns), t = tid / F
        int flp, t;
        if (a.col_direct) {
            t = (int)fdiv((uint32_t)tid, a.div_F);
            flp = tid - t * F;
        } else {
            flp = tid / T;
            t = tid - flp * T;
        }
        const bool active = flp < nf;
        const int lbase = (active ? flp : 0) * LP;
        const int fsel = active ? flp : 0;
        const float2 *gin = static_cast<const float2 *>(a.in) + ((a.direct_in || a.col_direct) ? in_base[fsel] : 0);
        float2 *gout = static_cast<float2 *>(a.out) + ((a.direct_out || a.col_direct) ? out_base[fsel] : 0);
        run(a, re, im, lbase, t, active, gin, gout, nf, in_base, out_base, col[fsel]);
    }
    if (a.direct_out || a.col_direct) return;

    // ---------------------------------------------------------------- store tile
    const int out_len = a.out_len;
    if (a.out_sep) {
        // split planes (C2C), with the four-step twiddle when the stage carries one
        float *dre = static_cast<float *>(a.out), *dim = a.out_im;
        for (int i = tid; i < F * out_len; i += nthreads) {
            int fl, j;
            if (a.tout.unit_f) {
                j = (int)fdiv((uint32_t)i, a.div_F);
                fl = i - j * F;
            } else {
                fl = (int)fdiv((uint32_t)i, a.div_out_len);
                j = i - fl * out_len;
            }
            if (fl < nf) {
                float2 v = make_float2(re[fl * LP + lds_pad(j)], im[fl * LP + lds_pad(j)]);
                if (a.tw4_ntot != 0) v = cmul(v, tw4_root(a, col[fl] * (int64_t)j));
                const int64_t o = out_base[fl] + (int64_t)j * a.tout.es;
                dre[o] = v.x;
                dim[o] = v.y;
            }
        }
    } else if (a.io_mode == IO_R2C_EVEN) {
        // X[k] = Z[k] (1 - i w^k)/2 + conj Z[L-k] (1 + i w^k)/2, k = 0..L (Z[L] = Z[0]). Unit kk > 0
        // of a transform writes bins kk and L - kk from the same two LDS values (w^{L-k} =
        // -conj w^k); unit 0 writes bins 0 and L. For even L the self-paired bin L/2 is one more
        // step per transform. A transform has (L+1)/2 units and the block F * T threads; all LDS
        // and root loads come before the first store.
        constexpr int H = L / 2;
        constexpr int NUN = (L + 1) / 2;
        constexpr bool EVEN = (L % 2) == 0;
        constexpr int NU = (NUN + T - 1) / T;
        float2 *dst = static_cast<float2 *>(a.out);
        const int64_t es = a.tout.es;
        float2 zk[NU], zm[NU], wk[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int i = tid + u * nthreads;
            const int fl = i / NUN;
            const int kk = i - fl * NUN;
            if (i < F * NUN && fl < nf) {
                const float *pr = re + fl * LP, *pi = im + fl * LP;
                const int km = kk == 0 ? 0 : L - kk;
                zk[u] = make_float2(pr[lds_pad(kk)], pi[lds_pad(kk)]);
                zm[u] = make_float2(pr[lds_pad(km)], pi[lds_pad(km)]);
                wk[u] = a.rtw[kk];
            }
        }
        if (EVEN && tid < nf) {
            const float2 zh = make_float2(re[tid * LP + lds_pad(H)], im[tid * LP + lds_pad(H)]);
            dst[out_base[tid] + (int64_t)H * es] = r2c_bin(zh, zh, a.rtw[H]);
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int i = tid + u * nthreads;
            const int fl = i / NUN;
            const int kk = i - fl * NUN;
            if (i >= F * NUN || fl >= nf) continue;
            float2 *o = dst + out_base[fl];
            const float2 w2 = make_float2(-wk[u].x, wk[u].y);  // w^{L - kk}
            o[(int64_t)kk * es] = r2c_bin(zk[u], zm[u], wk[u]);
            o[(int64_t)(L - kk) * es] = r2c_bin(zm[u], zk[u], w2);
        }
    } else if (a.io_mode == IO_C2R_FULL) {
        float *dst = static_cast<float *>(a.out);
        for (int i = tid; i < F * out_len; i += nthreads) {
            int fl, j;
            if (a.tout.unit_f) {
                j = (int)fdiv((uint32_t)i, a.div_F);
                fl = i - j * F;
            } else {
                fl = (int)fdiv((uint32_t)i, a.div_out_len);
                j = i - fl * out_len;
            }
            if (fl < nf) dst[out_base[fl] + (int64_t)j * a.tout.es] = re[fl * LP + lds_pad(j)];
        }
    } else if (a.tout.vec2 && a.tw4_ntot == 0) {
        constexpr int NV = ((L + 1) / 2 + T - 1) / T;
        float4 *dst = static_cast<float4 *>(a.out);
        const int half = out_len >> 1;
        const int total = F * half;
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int i = tid + u * nthreads;
            const int fl = (int)fdiv((uint32_t)i, a.div_out_vec);
            if (i < total && fl < nf) {
                const int jj = i - fl * half;
                const int j = 2 * jj;
                const float *pr = re + fl * LP, *pi = im + fl * LP;
                dst[(out_base[fl] >> 1) + jj] =
                    make_float4(pr[lds_pad(j)], pi[lds_pad(j)], pr[lds_pad(j + 1)], pi[lds_pad(j + 1)]);
            }
        }
    } else if (a.tout.unit_f && a.col_fixed) {
        constexpr int NU = (L + T - 1) / T;
        const int j0 = (int)fdiv((uint32_t)tid, a.div_F);
        const int fl = tid - j0 * F;
        if (fl < nf) {
            float2 *p = static_cast<float2 *>(a.out) + out_base[fl];
            const int64_t es = a.tout.es;
            const float *pr = re + fl * LP, *pi = im + fl * LP;
            float2 v[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const int j = j0 + u * T;
                if (j < out_len) v[u] = make_float2(pr[lds_pad(j)], pi[lds_pad(j)]);
            }
            if (a.tw4_ntot != 0) {
                // four-step twiddle: all table loads issued before the first multiply
                const int64_t c = col[fl];
                float2 w[NU];
#pragma unroll
                for (int u = 0; u < NU; u++) {
                    const int j = j0 + u * T;
                    if (j < out_len) w[u] = tw4_root(a, c * (int64_t)j);
                }
#pragma unroll
                for (int u = 0; u < NU; u++) v[u] = cmul(v[u], w[u]);
            }
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const int j = j0 + u * T;
                if (j < out_len) p[(int64_t)j * es] = v[u];
            }
        }
    } else {
        float2 *dst = static_cast<float2 *>(a.out);
        for (int i = tid; i < F * out_len; i += nthreads) {
            int fl, j;
            if (a.tout.unit_f) {
                j = (int)fdiv((uint32_t)i, a.div_F);
                fl = i - j * F;
            } else {
                fl = (int)fdiv((uint32_t)i, a.div_out_len);
                j = i - fl * out_len;
            }
            if (fl < nf) {
                float2 v = make_float2(re[fl * LP + lds_pad(j)], im[fl * LP + lds_pad(j)]);
                if (a.tw4_ntot != 0) v = cmul(v, tw4_root(a, col[fl] * (int64_t)j));
                dst[out_base[fl] + (int64_t)j * a.tout.es] = v;
            }
        }
    }
}

template <int L, int T, int S, int... Rs>
__global__ __launch_bounds__(1024) void stockham_kernel(const StockhamArgs a)
{
    extern __shared__ float lds[];
    stockham_body<L, T, S>(a, lds, StaticRunner<L, T, S, Rs...>{});
}

}  // namespace sip
