// Elementwise re-layout kernels used between Stockham stages: real <-> complex conversion,
// Hermitian extension / truncation, the split post- and pre-processing of even-length real
// transforms planned in several passes, and the three pointwise steps of Bluestein's algorithm.
#pragma once
#include <stdint.h>

#include "stockham_args.h"

namespace sip {

enum AuxMode : int32_t {
    AUX_REAL_TO_CPLX = 0,  // out[j] = (in_real[j], 0),                 j < n
    AUX_CPLX_TRUNC = 1,    // out[k] = in[k],                           k < out_len
    AUX_HERM_EXTEND = 2,   // out[j] = j <= n/2 ? in[j] : conj in[n-j]; imag of 0 and n/2 dropped
    AUX_CPLX_REAL = 3,     // out_real[j] = re in[j],                   j < n
    AUX_R2C_POST = 4,      // Z (n = L complex) -> X (L+1 bins) with roots tab0[k] = w^k, k <= L
    AUX_C2R_PRE = 5,       // X (L+1 bins) -> P (L complex) with roots tab0
    AUX_BLUE_PRE = 6,      // out[m] = m < n ? in[m] * tab0[m] : 0,     m < out_len (= M)
    AUX_BLUE_MUL = 7,      // out[m] = in[m] * tab0[m],                 m < out_len
    AUX_BLUE_POST = 8,     // out[k] = in[k] * tab0[k] * scale,         k < out_len (= n)
    AUX_ISTFT_OLA = 9,     // windowed overlap-add of n-point frames (hop, frames, start; see aux.hip)
};

struct AuxRows {
    int64_t n0;        // row r = r0 + n0 * r1
    int64_t s0, s1;    // strides of r0, r1 (elements)
    int64_t es;        // element stride
    FastDiv div_n0;
    // Gathered rows: when tab != null, the offset contributed by index r<tab_dim> is tab[r<tab_dim>]
    // (elements from the launch's base pointer) instead of r<tab_dim> * s<tab_dim>.
    const int64_t *tab;
    int32_t tab_dim;
    int32_t pad_;
};

struct AuxArgs {
    const void *in;
    void *out;
    const float *in_im;   // split-plane input (C2C modes): real parts at (float *)in
    float *out_im;        // split-plane output: real parts at (float *)out
    int32_t in_sep, out_sep;
    const float *win;     // ISTFT window (n floats)
    int64_t hop, frames, start;
    const float2 *tab0;
    int64_t rows;
    int64_t n;          // logical transform length
    int64_t out_len;    // elements written per row
    float scale;
    int32_t mode;
    AuxRows rin, rout;
};

}  // namespace sip
