// MFMA (split-precision) Stockham instances: lengths whose radix sequence is radix-16 stages plus
// at most one smaller radix, T = L / 16 threads per transform. One table per precision mode.
#include "mfma_dft.h"
#include "mfma_split.h"

namespace sip {

template <class... A>
constexpr int count_args(A...) { return (int)sizeof...(A); }

// fwd / inv: two-plane tiles (K = 16 instructions); rows_fwd / rows_inv: single-plane contiguous
// rows with K = 32 instructions; cols[0..1]:
// single-plane column tiles with K = 32 instructions (no four-step twiddle)
#define SIP_MF(L, DT, NT, ...)                                                                     \
    {L, L / 16, count_args(__VA_ARGS__), {__VA_ARGS__},                                              \
     (const void *)&stockham_mfma_kernel<L, L / 16, -1, DT, NT, __VA_ARGS__>,                       \
     (const void *)&stockham_mfma_kernel<L, L / 16, 1, DT, NT, __VA_ARGS__>,                        \
     (const void *)&stockham_mf2_rows_kernel<L, L / 16, -1, DT, NT, __VA_ARGS__>,                   \
     (const void *)&stockham_mf2_rows_kernel<L, L / 16, 1, DT, NT, __VA_ARGS__>,                    \
     {(const void *)&stockham_mf2_cols_kernel<L, L / 16, -1, DT, NT, __VA_ARGS__>,                  \
      (const void *)&stockham_mf2_cols_kernel<L, L / 16, 1, DT, NT, __VA_ARGS__>, nullptr, nullptr}, \
     {nullptr, nullptr, nullptr, nullptr}, {nullptr, nullptr}}

#define SIP_MF_SET(DT, NT)                                                                         \
    {SIP_MF(128, DT, NT, 16, 8), SIP_MF(256, DT, NT, 16, 16), SIP_MF(512, DT, NT, 16, 16, 2),       \
     SIP_MF(1024, DT, NT, 16, 16, 4), SIP_MF(2048, DT, NT, 16, 16, 8), SIP_MF(4096, DT, NT, 16, 16, 16)}

extern const StockhamInstance g_mfma_f16[6] = SIP_MF_SET(MF_F16, 2);
extern const StockhamInstance g_mfma_bf16[6] = SIP_MF_SET(MF_BF16, 2);
extern const StockhamInstance g_mfma_bf16x3[6] = SIP_MF_SET(MF_BF16, 3);

}  // namespace sip
