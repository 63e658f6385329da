// Launch descriptors shared by the host planner and the gfx950 Stockham kernels.
//
// One launch runs `nfft` independent transforms of length L. Each transform is staged in LDS
// (separate re / im float planes with one pad word every 16 words), goes through the radix passes
// of the Stockham autosort recursion (the radix sequence is a compile-time property of the kernel
// instance), and is written back. Transform f is addressed as f = f0 + n0 * f1 (f0 < n0); element j
// of it sits at  base + f0*s0 + f1*s1 + j*es  (in complex elements for complex data, in floats for
// real data). With these strides one kernel covers contiguous batches, the column passes of 2D/3D
// transforms and the passes of a four-step transform, always with coalesced global accesses: the
// tile is walked along whichever axis has unit stride.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sip {

// Exact unsigned division of n < 2^31 by a runtime constant d >= 1
// (round-up multiplier: q = (umulhi(n, m) + n) >> s, Granlund & Montgomery, PLDI 1994).
struct FastDiv {
    uint32_t d;
    uint32_t m;
    uint32_t s;
    uint32_t pad_;
};

enum IoMode : int32_t {
    IO_C2C = 0,        // complex in, complex out
    IO_R2C_EVEN = 1,   // real N = 2L in (read as L complex), L-point FFT, split post-process to L+1 bins
    IO_C2R_EVEN = 2,   // L+1 bins in, pre-process to L complex, L-point FFT, 2L reals out
    IO_R2C_FULL = 3,   // real L in, L-point FFT, L/2+1 bins out (odd lengths)
    IO_C2R_FULL = 4,   // L/2+1 bins in, Hermitian extension to L, L-point FFT, L reals out (odd)
};

struct TileDesc {
    int64_t n0, n1;    // f = f0 + n0 * (f1 + n1 * f2)
    int64_t s0, s1, s2;  // strides of f0, f1, f2
    int64_t es;        // stride between elements of one transform
    int32_t unit_f;    // 1 when consecutive transforms are adjacent (column pass): walk f fastest
    int32_t vec2;      // 1 when es == 1 and every row start is 16-byte aligned (float4 accesses)
    FastDiv div_n0, div_n1;
    // Gathered rows: when tab != null, the offset contributed by index f<tab_dim> is tab[f<tab_dim>]
    // (elements from the launch's base pointer) instead of f<tab_dim> * s<tab_dim>.
    const int64_t *tab;
    int32_t tab_dim;
    int32_t pad_;
};

struct StockhamArgs {
    const void *in;
    void *out;
    const float2 *tw;      // pass twiddles (FP64-generated, rounded to FP32), layout fixed by the radix sequence
    const float2 *rtw;     // real-transform roots w^k = e^{s 2 pi i k/(2L)}, k = 0..L
    const float2 *tw2;     // fused 2D tiles: pass twiddles of the column (X) transforms
    const float2 *tw4_lo;  // four-step twiddles W^m = lo[m & mask] * hi[m >> shift], W = e^{s 2 pi i/Ntot}
    const float2 *tw4_hi;
    const unsigned short *mfma_w;  // split DFT-matrix fragments of the MFMA radix-16 stages
    int64_t nfft;          // number of transforms in this launch (< 2^31)
    int64_t tw4_ntot;      // 0 = no four-step twiddle
    int64_t tw4_mask;
    int32_t tw4_shift;
    int32_t F;             // transforms per workgroup
    int32_t LP;            // LDS pitch (floats) of one transform in a plane
    int32_t io_mode;
    int32_t in_len;        // elements per transform on input (L, L+1 bins, L reals ...)
    int32_t out_len;       // elements per transform on output
    int32_t direct_in;     // first radix pass reads HBM directly (contiguous complex rows)
    int32_t direct_out;    // last radix pass writes HBM directly (contiguous complex rows)
    int32_t split_lds;     // launched as the single-plane row kernel (direct rows only)
    int32_t xcd_remap;     // give each XCD a contiguous range of tiles
    int32_t col_fixed;     // column tiles: block size is a multiple of F (one transform per thread)
    int32_t col_direct;    // column tile (C2C, strided in and out): lanes walk transforms, the first
                           // pass reads and the last pass writes HBM directly
    const float *in_im;    // split-plane input: real parts at (float *)in, imaginary parts here
    float *out_im;         // split-plane output: real parts at (float *)out, imaginary parts here
    int32_t in_sep;        // the input is two float planes (element offsets as for complex)
    int32_t out_sep;       // the output is two float planes
    int32_t nt_out;        // single-plane rows / matrix-core rows / column tiles without the
                           // four-step twiddle: the last pass stores nontemporal
    int32_t nt_in;         // single-plane rows / column tiles / rows-in columns-out: the first pass
                           // loads nontemporal
    FastDiv div_in_len;    // divides by in_len
    FastDiv div_out_len;   // divides by out_len
    FastDiv div_F;         // divides by F
    FastDiv div_col;       // four-step column index = f0(in) / col_div
    FastDiv div_in_vec;    // divides by in_len / 2 (16-byte path)
    FastDiv div_out_vec;   // divides by out_len / 2
    TileDesc tin, tout;
};

// Precision modes of the radix-16 stages (plain = FP32 VALU butterflies).
enum PrecisionMode : int32_t {
    PREC_FP32 = 0,       // plain form
    PREC_F16_SPLIT = 1,  // MFMA, FP16 hi+lo operands, per-column power-of-two scaling
    PREC_BF16_SPLIT = 2, // MFMA, BF16 hi+lo operands
    PREC_BF16_SPLIT3 = 3 // MFMA, BF16 hi+mid+lo operands
};

// Host-side view of one compiled kernel instance.
typedef void (*StockhamKernelFn)(StockhamArgs);
struct StockhamInstance {
    int32_t L;
    int32_t T;               // threads per transform
    int32_t npass;
    int32_t radix[12];
    const void *fwd;         // __global__ function handles (hipLaunchKernel symbol)
    const void *inv;
    const void *rows_fwd;    // single-plane contiguous-row kernels (T >= 16 only, else null)
    const void *rows_inv;
    const void *cols[4];     // single-plane column-tile kernels, fwd / inv without and fwd / inv
                             // with the four-step twiddle (L <= 1024 only, else null)
    const void *real_rows[4];  // single-plane R2C fwd / inv, C2R fwd / inv rows (T >= 32; L <= 8192
                               // and the longer L where they measured faster; else null)
    const void *rc[2];         // single-plane rows-in / columns-out fwd / inv (four-step second
                               // stage; L <= 1024 with two or more passes and T >= 8, else null)
};

// Registry entry of a fused 2D tile instance (stockham2d.h). The radix sequences, and so the
// twiddle tables, are those of the 1D leaf instances of lengths X and Y.
struct Stockham2dInstance {
    int32_t X, TX, Y, TY;
    int32_t threads;
    int32_t lds_bytes;
    const void *fwd;
    const void *inv;
};

}  // namespace sip
