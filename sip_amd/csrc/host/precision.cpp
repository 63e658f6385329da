// Precision form: lookup of the MFMA (split FP16 / BF16) Stockham instances and the split
// DFT-matrix operand tables of their radix-16 stages (built in FP64 on the host, split into
// FP16 / BF16 terms with round-to-nearest-even, laid out in the MFMA A-operand lane order).
#include <cmath>
#include <cstring>

#include "plan.h"

namespace sip {

extern const StockhamInstance g_mfma_f16[6];
extern const StockhamInstance g_mfma_bf16[6];
extern const StockhamInstance g_mfma_bf16x3[6];

const StockhamInstance *find_mfma_instance(int64_t L, int mode)
{
    const StockhamInstance *tab = mode == PREC_F16_SPLIT ? g_mfma_f16
                                  : mode == PREC_BF16_SPLIT ? g_mfma_bf16
                                  : mode == PREC_BF16_SPLIT3 ? g_mfma_bf16x3 : nullptr;
    if (!tab) return nullptr;
    for (int i = 0; i < 6; i++)
        if (tab[i].L == L) return &tab[i];
    return nullptr;
}

// IEEE binary16 bits of x, round to nearest even (subnormals kept, overflow to infinity).
static uint16_t half_bits(double xd)
{
    const float x = (float)xd;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const int32_t exp = (int32_t)((u >> 23) & 0xFFu) - 127;
    uint32_t man = u & 0x7FFFFFu;
    if (exp > 15) return (uint16_t)(sign | 0x7C00u);
    if (exp >= -14) {
        uint32_t h = ((uint32_t)(exp + 15) << 10) | (man >> 13);
        const uint32_t rest = man & 0x1FFFu;
        if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    if (exp < -25) return (uint16_t)sign;
    man |= 0x800000u;
    const int shift = -14 - exp + 13;  // 14..25
    uint32_t h = man >> shift;
    const uint32_t rest = man & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
    if (rest > halfway || (rest == halfway && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

static double half_value(uint16_t h)
{
    const int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    double v = e == 0 ? std::ldexp((double)m, -24) : std::ldexp(1024.0 + m, e - 25);
    return s ? -v : v;
}

static uint16_t bf16_bits(double xd)
{
    const float x = (float)xd;
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

static double bf16_value(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

TableRef mfma_dft_table(int mode, int sign)
{
    TableRef t;
    t.key = "mfw:" + std::to_string(mode) + ":" + std::to_string(sign);
    t.gen = [mode, sign]() {
        const int nt = mode == PREC_BF16_SPLIT3 ? 3 : 2;
        const bool f16 = mode == PREC_F16_SPLIT;
        std::vector<uint16_t> v((size_t)nt * 2 * 2 * 64 * 4);
        const long double two_pi = 6.283185307179586476925286766559005768L;
        for (int rb = 0; rb < 2; rb++)
            for (int kb = 0; kb < 2; kb++)
                for (int lane = 0; lane < 64; lane++)
                    for (int t4 = 0; t4 < 4; t4++) {
                        const int q = lane % 16, r = 4 * (lane / 16) + t4;
                        const long double a = two_pi * (long double)((q * r) % 16) / 16.0L;
                        const double wr = (double)cosl(a), wi = (double)(sign * sinl(a));
                        // [[W_re, -W_im], [W_im, W_re]]
                        const double w = (rb == 0 && kb == 0) || (rb == 1 && kb == 1) ? wr : (rb == 0 ? -wi : wi);
                        double rest = w;
                        for (int tm = 0; tm < nt; tm++) {
                            const uint16_t b = f16 ? half_bits(rest) : bf16_bits(rest);
                            v[((((size_t)tm * 2 + rb) * 2 + kb) * 64 + lane) * 4 + t4] = b;
                            rest -= f16 ? half_value(b) : bf16_value(b);
                        }
                    }
        std::vector<float2> out((v.size() + 3) / 4);
        std::memcpy(out.data(), v.data(), v.size() * 2);
        return out;
    };
    return t;
}

// Operands of the K = 32 radix-16 passes (kernels/mfma_split.h): [term][part][lane][8] 16-bit
// values, part 0 = A_r (row q, column 2r: Re W[q][r], column 2r + 1: -Im W[q][r]), part 1 = A_i
// (Im W[q][r], Re W[q][r]); lane l holds row l % 16, columns 8 (l / 16) .. + 7.
TableRef mfma2_dft_table(int mode, int sign)
{
    TableRef t;
    t.key = "mfw2:" + std::to_string(mode) + ":" + std::to_string(sign);
    t.gen = [mode, sign]() {
        const int nt = mode == PREC_BF16_SPLIT3 ? 3 : 2;
        const bool f16 = mode == PREC_F16_SPLIT;
        std::vector<uint16_t> v((size_t)nt * 2 * 64 * 8);
        const long double two_pi = 6.283185307179586476925286766559005768L;
        for (int part = 0; part < 2; part++)
            for (int lane = 0; lane < 64; lane++)
                for (int k8 = 0; k8 < 8; k8++) {
                    const int q = lane % 16, col = 8 * (lane / 16) + k8, r = col / 2, c = col % 2;
                    const long double a = two_pi * (long double)((q * r) % 16) / 16.0L;
                    const double wr = (double)cosl(a), wi = (double)(sign * sinl(a));
                    const double w = part == 0 ? (c == 0 ? wr : -wi) : (c == 0 ? wi : wr);
                    double rest = w;
                    for (int tm = 0; tm < nt; tm++) {
                        const uint16_t b = f16 ? half_bits(rest) : bf16_bits(rest);
                        v[(((size_t)tm * 2 + part) * 64 + lane) * 8 + k8] = b;
                        rest -= f16 ? half_value(b) : bf16_value(b);
                    }
                }
        std::vector<float2> out((v.size() + 3) / 4);
        std::memcpy(out.data(), v.data(), v.size() * 2);
        return out;
    };
    return t;
}

}  // namespace sip
