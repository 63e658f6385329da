// Small DFT kernels on registers for the Stockham passes (gfx950, FP32).
//
// dft<R, S>(v) overwrites v[0..R-1] with V[q] = sum_r v[r] e^{S 2 pi i r q / R}; S = -1 forward,
// S = +1 inverse. Radix 4/8/16 are split-radix style decompositions into 2- and 4-point DFTs
// with the internal roots written as FP32 roundings of their FP64 values; 3 and 5 use the
// conjugate-pair (Winograd-style) form; 7, 11 and 13 use the same pairing with constant tables;
// 6, 9, 10, 12, 15, 20 and 25 are two-factor Cooley-Tukey products of those with compile-time
// roots (double precision, rounded once).
#pragma once
#include <hip/hip_runtime.h>

namespace sip {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// multiply by S*i
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a)
{
    return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// multiply by e^{S i pi/4} = (1 + S i)/sqrt(2)
template <int S>
__device__ __forceinline__ float2 mul_w8_1(float2 a)
{
    const float h = 7.071067691e-01f;
    return S > 0 ? make_float2((a.x - a.y) * h, (a.x + a.y) * h) : make_float2((a.x + a.y) * h, (a.y - a.x) * h);
}

// multiply by e^{S i 3pi/4} = (-1 + S i)/sqrt(2)
template <int S>
__device__ __forceinline__ float2 mul_w8_3(float2 a)
{
    const float h = 7.071067691e-01f;
    return S > 0 ? make_float2((-a.x - a.y) * h, (a.x - a.y) * h) : make_float2((a.y - a.x) * h, (-a.x - a.y) * h);
}

template <int S>
__device__ __forceinline__ void dft2(float2 &a, float2 &b)
{
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int S>
__device__ __forceinline__ void dft4(float2 &x0, float2 &x1, float2 &x2, float2 &x3)
{
    float2 a = cadd(x0, x2), b = csub(x0, x2);
    float2 c = cadd(x1, x3), d = mul_si<S>(csub(x1, x3));
    x0 = cadd(a, c);
    x2 = csub(a, c);
    x1 = cadd(b, d);
    x3 = csub(b, d);
}

template <int S>
__device__ __forceinline__ void dft3(float2 &x0, float2 &x1, float2 &x2)
{
    const float c = -0.5f, s = 8.660253882e-01f;
    float2 t1 = cadd(x1, x2), t2 = csub(x1, x2);
    float2 a = make_float2(fmaf(c, t1.x, x0.x), fmaf(c, t1.y, x0.y));
    float2 b = mul_si<S>(cscale(t2, s));
    x0 = cadd(x0, t1);
    x1 = cadd(a, b);
    x2 = csub(a, b);
}

template <int S>
__device__ __forceinline__ void dft5(float2 &x0, float2 &x1, float2 &x2, float2 &x3, float2 &x4)
{
    const float c1 = 3.090170026e-01f, c2 = -8.090170026e-01f;
    const float s1 = 9.510565400e-01f, s2 = 5.877852440e-01f;
    float2 t1 = cadd(x1, x4), t2 = cadd(x2, x3), t3 = csub(x1, x4), t4 = csub(x2, x3);
    float2 a1 = make_float2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
    float2 a2 = make_float2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
    float2 b1 = mul_si<S>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    float2 b2 = mul_si<S>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    x0 = cadd(x0, cadd(t1, t2));
    x1 = cadd(a1, b1);
    x4 = csub(a1, b1);
    x2 = cadd(a2, b2);
    x3 = csub(a2, b2);
}

// Odd prime P with cos/sin(2 pi r/P), r = 1..(P-1)/2, in tables.
template <int P>
struct PrimeRoots;
template <>
struct PrimeRoots<7> {
    static __device__ __forceinline__ float c(int r)
    {
        return r == 1 ? 6.234897971e-01f : r == 2 ? -2.225209326e-01f : -9.009688497e-01f;
    }
    static __device__ __forceinline__ float s(int r)
    {
        return r == 1 ? 7.818315029e-01f : r == 2 ? 9.749279022e-01f : 4.338837266e-01f;
    }
};
template <>
struct PrimeRoots<11> {
    static __device__ __forceinline__ float c(int r)
    {
        constexpr float t[5] = {8.412535191e-01f, 4.154150188e-01f, -1.423148364e-01f, -6.548607349e-01f,
                                -9.594929814e-01f};
        return t[r - 1];
    }
    static __device__ __forceinline__ float s(int r)
    {
        constexpr float t[5] = {5.406408310e-01f, 9.096319675e-01f, 9.898214340e-01f, 7.557495832e-01f,
                                2.817325592e-01f};
        return t[r - 1];
    }
};
template <>
struct PrimeRoots<13> {
    static __device__ __forceinline__ float c(int r)
    {
        constexpr float t[6] = {8.854560256e-01f, 5.680647492e-01f, 1.205366775e-01f, -3.546048999e-01f,
                                -7.485107780e-01f, -9.709418416e-01f};
        return t[r - 1];
    }
    static __device__ __forceinline__ float s(int r)
    {
        constexpr float t[6] = {4.647231698e-01f, 8.229838610e-01f, 9.927088618e-01f, 9.350162148e-01f,
                                6.631226540e-01f, 2.393156588e-01f};
        return t[r - 1];
    }
};

template <int P, int S>
__device__ __forceinline__ void dft_prime(float2 *v)
{
    constexpr int H = (P - 1) / 2;
    float2 t[H + 1], u[H + 1];
#pragma unroll
    for (int r = 1; r <= H; r++) {
        t[r] = cadd(v[r], v[P - r]);
        u[r] = csub(v[r], v[P - r]);
    }
    float2 x0 = v[0];
    float2 sum = x0;
#pragma unroll
    for (int r = 1; r <= H; r++) sum = cadd(sum, t[r]);
#pragma unroll
    for (int q = 1; q <= H; q++) {
        float2 a = x0, b = make_float2(0.f, 0.f);
#pragma unroll
        for (int r = 1; r <= H; r++) {
            const int m = (r * q) % P;
            // cos(2 pi m/P) and sin(2 pi m/P) from the first-half tables by symmetry
            const float cm = (m <= H) ? PrimeRoots<P>::c(m) : PrimeRoots<P>::c(P - m);
            const float sm = (m <= H) ? PrimeRoots<P>::s(m) : -PrimeRoots<P>::s(P - m);
            a.x = fmaf(cm, t[r].x, a.x);
            a.y = fmaf(cm, t[r].y, a.y);
            b.x = fmaf(sm, u[r].x, b.x);
            b.y = fmaf(sm, u[r].y, b.y);
        }
        float2 ib = mul_si<S>(b);
        v[q] = cadd(a, ib);
        v[P - q] = csub(a, ib);
    }
    v[0] = sum;
}

template <int S>
__device__ __forceinline__ void dft8(float2 *v)
{
    // DIF radix-2 across stride 4, then two 4-point DFTs (even and odd outputs).
    float2 a0 = cadd(v[0], v[4]), b0 = csub(v[0], v[4]);
    float2 a1 = cadd(v[1], v[5]), b1 = mul_w8_1<S>(csub(v[1], v[5]));
    float2 a2 = cadd(v[2], v[6]), b2 = mul_si<S>(csub(v[2], v[6]));
    float2 a3 = cadd(v[3], v[7]), b3 = mul_w8_3<S>(csub(v[3], v[7]));
    dft4<S>(a0, a1, a2, a3);
    dft4<S>(b0, b1, b2, b3);
    v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
    v[1] = b0; v[3] = b1; v[5] = b2; v[7] = b3;
}

template <int S>
__device__ __forceinline__ void dft16(float2 *v)
{
    // 4 x 4: y[r][q] = DFT4_t(v[r + 4t])[q]; y[r][q] *= w16^{rq}; V[q + 4m] = DFT4_r(y[r][q])[m].
    const float c1 = 9.238795042e-01f, s1 = 3.826834261e-01f;
    float2 y[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float2 a = v[r], b = v[r + 4], c = v[r + 8], d = v[r + 12];
        dft4<S>(a, b, c, d);
        y[r][0] = a; y[r][1] = b; y[r][2] = c; y[r][3] = d;
    }
    // w16^1 = c1 + S i s1, w16^2 = (1 + S i)/sqrt2, w16^3 = s1 + S i c1,
    // w16^4 = S i, w16^6 = (-1 + S i)/sqrt2, w16^9 = -c1 - S i s1
    const float2 w1 = make_float2(c1, S * s1), w3 = make_float2(s1, S * c1), w9 = make_float2(-c1, -S * s1);
    y[1][1] = cmul(y[1][1], w1);
    y[1][2] = mul_w8_1<S>(y[1][2]);
    y[1][3] = cmul(y[1][3], w3);
    y[2][1] = mul_w8_1<S>(y[2][1]);
    y[2][2] = mul_si<S>(y[2][2]);
    y[2][3] = mul_w8_3<S>(y[2][3]);
    y[3][1] = cmul(y[3][1], w3);
    y[3][2] = mul_w8_3<S>(y[3][2]);
    y[3][3] = cmul(y[3][3], w9);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float2 a = y[0][q], b = y[1][q], c = y[2][q], d = y[3][q];
        dft4<S>(a, b, c, d);
        v[q] = a; v[q + 4] = b; v[q + 8] = c; v[q + 12] = d;
    }
}

// cos / sin (2 pi m / n), evaluated at compile time in double (Taylor series on [0, pi/4] after
// quarter-turn and octant reduction; quarter points exact) and rounded once to FP32.
constexpr double kCePi = 3.14159265358979323846;
constexpr double ce_sin_poly(double x)
{
    double x2 = x * x, term = x, sum = x;
    for (int k = 1; k < 14; k++) {
        term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ce_cos_poly(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; k++) {
        term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
constexpr void ce_cossin2pi(int m, int n, double &c, double &s)
{
    m %= n;
    const int q = (4 * m) / n;          // quarter turn
    const int rem = 4 * m - q * n;      // remainder in units of 1/(4n) turns
    double cc, ss;
    if (rem == 0) {
        cc = 1.0;
        ss = 0.0;
    } else {
        const double x = kCePi / 2.0 * (double)rem / (double)n;  // in (0, pi/2)
        if (2 * rem <= n) {
            cc = ce_cos_poly(x);
            ss = ce_sin_poly(x);
        } else {
            const double y = kCePi / 2.0 * (double)(n - rem) / (double)n;
            cc = ce_sin_poly(y);
            ss = ce_cos_poly(y);
        }
    }
    for (int k = 0; k < q; k++) {  // rotate by q quarter turns
        const double t = cc;
        cc = -ss;
        ss = t;
    }
    c = cc;
    s = ss;
}

template <int R>
struct RootTable {
    float c[R], s[R];
    constexpr RootTable() : c(), s()
    {
        for (int m = 0; m < R; m++) {
            double cc = 0, ss = 0;
            ce_cossin2pi(m, R, cc, ss);
            c[m] = (float)cc;
            s[m] = (float)ss;
        }
    }
};
template <int R>
constexpr RootTable<R> kRoots = RootTable<R>();

template <int R, int S>
__device__ __forceinline__ void dft(float2 *v);

// Composite radix R = P * Q in registers (Cooley-Tukey with the internal roots as constants):
// y[n2][k1] = DFT_P over n1 of v[Q n1 + n2], y[n2][k1] *= W_R^{n2 k1}, V[k1 + P k2] = DFT_Q over n2.
template <int P, int Q, int S>
__device__ __forceinline__ void dft_pq(float2 *v)
{
    constexpr int R = P * Q;
    float2 y[Q][P];
#pragma unroll
    for (int n2 = 0; n2 < Q; n2++) {
        float2 t[P];
#pragma unroll
        for (int n1 = 0; n1 < P; n1++) t[n1] = v[Q * n1 + n2];
        dft<P, S>(t);
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) y[n2][k1] = t[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < Q; n2++)
#pragma unroll
        for (int k1 = 1; k1 < P; k1++) {
            const int m = (n2 * k1) % R;
            y[n2][k1] = cmul(y[n2][k1], make_float2(kRoots<R>.c[m], S * kRoots<R>.s[m]));
        }
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) {
        float2 t[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) t[n2] = y[n2][k1];
        dft<Q, S>(t);
#pragma unroll
        for (int k2 = 0; k2 < Q; k2++) v[k1 + P * k2] = t[k2];
    }
}

template <int R, int S>
__device__ __forceinline__ void dft(float2 *v)
{
    if constexpr (R == 2) {
        dft2<S>(v[0], v[1]);
    } else if constexpr (R == 3) {
        dft3<S>(v[0], v[1], v[2]);
    } else if constexpr (R == 4) {
        dft4<S>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 5) {
        dft5<S>(v[0], v[1], v[2], v[3], v[4]);
    } else if constexpr (R == 8) {
        dft8<S>(v);
    } else if constexpr (R == 16) {
        dft16<S>(v);
    } else if constexpr (R == 6) {
        dft_pq<2, 3, S>(v);
    } else if constexpr (R == 9) {
        dft_pq<3, 3, S>(v);
    } else if constexpr (R == 10) {
        dft_pq<2, 5, S>(v);
    } else if constexpr (R == 12) {
        dft_pq<4, 3, S>(v);
    } else if constexpr (R == 15) {
        dft_pq<3, 5, S>(v);
    } else if constexpr (R == 20) {
        dft_pq<4, 5, S>(v);
    } else if constexpr (R == 25) {
        dft_pq<5, 5, S>(v);
    } else {
        dft_prime<R, S>(v);
    }
}

}  // namespace sip
