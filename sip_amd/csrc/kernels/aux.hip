// Elementwise stages between Stockham passes (see aux_args.h). Each lane produces one output
// element; rows are addressed through (n0, s0, s1, es) so strided axes of 2D/3D arrays work too.
// These stages are HBM-bound copies with at most two complex multiplies per element.
#include <hip/hip_runtime.h>

#include "aux_args.h"
#include "butterflies.h"

namespace sip {

__device__ __forceinline__ uint32_t fdiv_aux(uint32_t n, const FastDiv &d) { return (__umulhi(n, d.m) + n) >> d.s; }

__device__ __forceinline__ int64_t row_base(uint32_t r, const AuxRows &d)
{
    const uint32_t r1 = fdiv_aux(r, d.div_n0);
    const uint32_t r0 = r - r1 * (uint32_t)d.n0;
    int64_t o0 = (int64_t)r0 * d.s0, o1 = (int64_t)r1 * d.s1;
    if (d.tab) {
        if (d.tab_dim == 0) o0 = d.tab[r0];
        else o1 = d.tab[r1];
    }
    return o0 + o1;
}

// Complex element at offset o of the input / output, interleaved or split into two planes.
__device__ __forceinline__ float2 cin_at(const AuxArgs &a, int64_t o)
{
    if (a.in_sep) return make_float2(static_cast<const float *>(a.in)[o], a.in_im[o]);
    return static_cast<const float2 *>(a.in)[o];
}
__device__ __forceinline__ void cout_at(const AuxArgs &a, int64_t o, float2 v)
{
    if (a.out_sep) {
        static_cast<float *>(a.out)[o] = v.x;
        a.out_im[o] = v.y;
    } else {
        static_cast<float2 *>(a.out)[o] = v;
    }
}

__global__ __launch_bounds__(256) void aux_kernel(const AuxArgs a)
{
    const int64_t total = a.rows * a.out_len;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t r = i / a.out_len;
        const int64_t k = i - r * a.out_len;
        const int64_t ib = row_base((uint32_t)r, a.rin);
        const int64_t ob = row_base((uint32_t)r, a.rout);
        const float2 *cin = static_cast<const float2 *>(a.in);
        float2 *cout = static_cast<float2 *>(a.out);
        switch (a.mode) {
            case AUX_REAL_TO_CPLX: {
                const float *rin = static_cast<const float *>(a.in);
                cout[ob + k * a.rout.es] = make_float2(rin[ib + k * a.rin.es], 0.f);
                break;
            }
            case AUX_CPLX_TRUNC:
                cout_at(a, ob + k * a.rout.es, cin_at(a, ib + k * a.rin.es));
                break;
            case AUX_HERM_EXTEND: {
                const int64_t h = a.n / 2;
                float2 v = (k <= h) ? cin[ib + k * a.rin.es] : cconj(cin[ib + (a.n - k) * a.rin.es]);
                if (k == 0 || (2 * k == a.n)) v.y = 0.f;
                cout[ob + k * a.rout.es] = v;
                break;
            }
            case AUX_CPLX_REAL: {
                float *rout = static_cast<float *>(a.out);
                rout[ob + k * a.rout.es] = cin[ib + k * a.rin.es].x;
                break;
            }
            case AUX_R2C_POST: {
                // X[k] = Z[k] (1 - i w^k)/2 + conj Z[L-k] (1 + i w^k)/2
                const int64_t L = a.n;
                const int64_t ka = (k == L) ? 0 : k;
                const int64_t kb = (k == 0) ? 0 : L - k;
                const float2 z = cin[ib + ka * a.rin.es];
                const float2 zc = cconj(cin[ib + kb * a.rin.es]);
                const float2 iw = mul_si<1>(a.tab0[k]);
                const float2 A = make_float2(0.5f * (1.f - iw.x), -0.5f * iw.y);
                const float2 B = make_float2(0.5f * (1.f + iw.x), 0.5f * iw.y);
                cout[ob + k * a.rout.es] = cadd(cmul(z, A), cmul(zc, B));
                break;
            }
            case AUX_C2R_PRE: {
                // P[k] = (X[k] + conj X[L-k]) + i w^k (X[k] - conj X[L-k]); imag of X[0], X[L] dropped
                const int64_t L = a.n;
                float2 xk = cin[ib + k * a.rin.es];
                float2 xm = cin[ib + (L - k) * a.rin.es];
                if (k == 0) {
                    xk.y = 0.f;
                    xm.y = 0.f;
                }
                const float2 c = cconj(xm);
                const float2 s = cadd(xk, c), d = csub(xk, c);
                cout[ob + k * a.rout.es] = cadd(s, mul_si<1>(cmul(a.tab0[k], d)));
                break;
            }
            case AUX_BLUE_PRE:
                cout[ob + k * a.rout.es] = (k < a.n) ? cmul(cin_at(a, ib + k * a.rin.es), a.tab0[k]) : make_float2(0.f, 0.f);
                break;
            case AUX_BLUE_MUL:
                cout[ob + k * a.rout.es] = cmul(cin[ib + k * a.rin.es], a.tab0[k]);
                break;
            case AUX_ISTFT_OLA: {
                // Output sample k of channel r is m = k + start of the overlap-add span; the frames
                // t with t*hop <= m < t*hop + n cover it. y = sum_t frame_t[m - t hop] w[m - t hop],
                // env = sum_t w[m - t hop]^2, out = y / (n env) (the 1/n of the inverse DFT).
                const int64_t n = a.n, hop = a.hop;
                const int64_t m = k + a.start;
                int64_t t_lo = m >= n ? (m - n) / hop + 1 : 0;
                int64_t t_hi = m / hop;
                if (t_hi > a.frames - 1) t_hi = a.frames - 1;
                float2 acc = make_float2(0.f, 0.f);
                float env = 0.f;
                for (int64_t t = t_lo; t <= t_hi; t++) {
                    const int64_t j = m - t * hop;
                    const float w = a.win[j];
                    const float2 f = cin[ib + t * n + j];
                    acc.x = fmaf(f.x, w, acc.x);
                    acc.y = fmaf(f.y, w, acc.y);
                    env = fmaf(w, w, env);
                }
                const float d = (float)n * env;
                cout[ob + k * a.rout.es] = make_float2(acc.x / d, acc.y / d);
                break;
            }
            case AUX_BLUE_POST:
                cout_at(a, ob + k * a.rout.es, cscale(cmul(cin[ib + k * a.rin.es], a.tab0[k]), a.scale));
                break;
            default:
                break;
        }
    }
}

hipError_t aux_launch(const AuxArgs &a, hipStream_t stream)
{
    const int64_t total = a.rows * a.out_len;
    if (total <= 0) return hipSuccess;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(aux_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace sip
