/*
 * FFT oracle: CPU restatement of the transforms computed by the reference's FFT path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links or calls this file; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it, as the checker.
 *
 * What it restates (reference semantics, not its kernels):
 *  - C2C: X[k] = sum_n x[n] e^{s 2 pi i k n / N}, s = -1 forward, +1 inverse, inverse unnormalised
 *    (include/fft_api.h:36-39; no 1/N scaling anywhere on the path, core/fft/fft_api.cpp:1001-1045;
 *    rotation matrix convention core/fft/fftcore/dft_core.cpp:63-112).
 *  - R2C: bins 0..N/2 of the transform of a real sequence, with the sign chosen by the direction
 *    (core/fft/fftcore/dft_r2c_core.cpp:96-101: sin term scaled by forward ? -1 : +1).
 *  - C2R: real output of the transform of the Hermitian extension of N/2+1 input bins; the imaginary
 *    parts of bin 0 (and bin N/2 for even N) are ignored because they multiply sin(0) / sin(pi n)
 *    (core/fft/fftcore/dft_c2r_core.cpp:90-105 folds bins i > N/2 onto bin N-i with conjugation;
 *    direction picks the sign of the sin rows).
 *  - Multi-dimensional transforms are separable: 2D/3D C2C apply the 1D transform along every
 *    axis (core/fft/fft_api.cpp:527-697). R2C runs the real transform on the last axis first, then
 *    C2C on the others (fft_api.cpp:559-576); C2R runs C2C on the leading axes, then C2R on the last
 *    (fft_api.cpp:538-558).
 *
 * Algorithms: an FP64 Stockham autosort FFT over the prime factorisation of N (any N), with all
 * roots of unity taken from one table w[m] = e^{s 2 pi i m/N} computed in long double and rounded
 * once; and an O(N^2) long-double brute-force DFT used to cross-check the FFT. An FP32 copy of the
 * same staged algorithm (twiddles rounded from FP64) serves as the CPU baseline.
 *
 * Seeded inputs: PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
 * Algorithms for Random Number Generation", 2014; XSH-RR output, 64-bit LCG state); uniform values
 * (u32 >> 8) * 2^-24 * 2 - 1; Gaussian values by Box-Muller in FP64, then rounded to FP32.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cd;
typedef struct { float re, im; } cf;

/* ---------------------------------------------------------------- PCG32 seeded generators */

typedef struct { uint64_t state, inc; } pcg32_t;

static uint32_t pcg32_next(pcg32_t *g)
{
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ULL + g->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

static void pcg32_seed(pcg32_t *g, uint64_t seed, uint64_t stream)
{
    g->state = 0u;
    g->inc = (stream << 1u) | 1u;
    pcg32_next(g);
    g->state += seed;
    pcg32_next(g);
}

/* Fill `count` floats with uniform[-1,1) values; stream id lets callers split work. */
void oracle_fill_uniform(float *dst, int64_t count, uint64_t seed, uint64_t stream)
{
    pcg32_t g;
    pcg32_seed(&g, seed, stream);
    for (int64_t i = 0; i < count; i++) {
        uint32_t u = pcg32_next(&g) >> 8;
        dst[i] = (float)((double)u * (1.0 / 16777216.0) * 2.0 - 1.0);
    }
}

/* Fill `count` floats with N(0,1) values (Box-Muller in FP64, rounded once). */
void oracle_fill_gaussian(float *dst, int64_t count, uint64_t seed, uint64_t stream)
{
    pcg32_t g;
    pcg32_seed(&g, seed, stream);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < count; i += 2) {
        double u1 = ((double)(pcg32_next(&g) >> 8) + 1.0) * (1.0 / 16777217.0); /* (0,1] */
        double u2 = (double)(pcg32_next(&g) >> 8) * (1.0 / 16777216.0);
        double r = sqrt(-2.0 * log(u1));
        dst[i] = (float)(r * cos(two_pi * u2));
        if (i + 1 < count) dst[i + 1] = (float)(r * sin(two_pi * u2));
    }
}

/* ---------------------------------------------------------------- factorisation and roots */

static int factorize(int64_t n, int64_t *f)
{
    /* Prime factors, 4s first (radix-4 passes keep the oracle fast), then ascending primes. */
    int m = 0;
    while (n % 4 == 0) { f[m++] = 4; n /= 4; }
    for (int64_t p = 2; p * p <= n; p++) {
        while (n % p == 0) { f[m++] = p; n /= p; }
    }
    if (n > 1) f[m++] = n;
    return m;
}

/* w[m] = exp(sign * 2 pi i m / n), computed in long double with the argument reduced to the first
 * octant so that every entry is the correctly rounded FP64 value of the exact root (to ~1 ulp). */
static void roots_ld(cd *w, int64_t n, int sign)
{
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int64_t m = 0; m < n; m++) {
        long double a = two_pi * (long double)m / (long double)n;
        long double c = cosl(a), s = sinl(a);
        /* exact values at the quarter points */
        if ((4 * m) % n == 0) {
            int64_t q = (4 * m) / n;
            c = (q == 0) ? 1.0L : (q == 2) ? -1.0L : 0.0L;
            s = (q == 1) ? 1.0L : (q == 3) ? -1.0L : 0.0L;
        }
        w[m].re = (double)c;
        w[m].im = (double)(sign * s);
    }
}

/* ---------------------------------------------------------------- FP64 Stockham FFT */

static void stockham_d(cd *x, cd *y, int64_t n, const int64_t *f, int nf, const cd *w, cd **result)
{
    int64_t ns = 1;
    cd *src = x, *dst = y;
    cd vbuf[64];
    for (int s = 0; s < nf; s++) {
        int64_t R = f[s];
        int64_t m = n / R;
        cd *v = (R <= 64) ? vbuf : (cd *)malloc(sizeof(cd) * (size_t)R);
        cd *V = (cd *)malloc(sizeof(cd) * (size_t)R);
        int64_t tstep = n / (ns * R);
        for (int64_t j = 0; j < m; j++) {
            int64_t k = j % ns;
            for (int64_t r = 0; r < R; r++) {
                cd a = src[j + r * m];
                cd t = w[(k * r * tstep) % n];
                v[r].re = a.re * t.re - a.im * t.im;
                v[r].im = a.re * t.im + a.im * t.re;
            }
            for (int64_t q = 0; q < R; q++) {
                double re = 0.0, im = 0.0;
                for (int64_t r = 0; r < R; r++) {
                    cd t = w[((r * q) % R) * m];
                    re += v[r].re * t.re - v[r].im * t.im;
                    im += v[r].re * t.im + v[r].im * t.re;
                }
                V[q].re = re;
                V[q].im = im;
            }
            int64_t d = (j / ns) * ns * R + k;
            for (int64_t q = 0; q < R; q++) dst[d + q * ns] = V[q];
        }
        if (v != vbuf) free(v);
        free(V);
        cd *t = src; src = dst; dst = t;
        ns *= R;
    }
    *result = src;
}

/* In-place batched FP64 C2C of `batch` contiguous length-n transforms (interleaved re,im).
 * sign = -1 forward, +1 inverse (unnormalised). Returns 0 on success. */
int oracle_c2c_f64(double *data, int64_t n, int64_t batch, int sign)
{
    if (n <= 0 || batch < 0) return -1;
    if (n == 1) return 0;
    int64_t f[64];
    int nf = factorize(n, f);
    cd *w = (cd *)malloc(sizeof(cd) * (size_t)n);
    if (!w) return -2;
    roots_ld(w, n, sign);
    int err = 0;
#pragma omp parallel
    {
        cd *x = (cd *)malloc(sizeof(cd) * (size_t)n);
        cd *y = (cd *)malloc(sizeof(cd) * (size_t)n);
        if (!x || !y) {
#pragma omp atomic write
            err = -2;
        } else {
#pragma omp for schedule(static)
            for (int64_t b = 0; b < batch; b++) {
                cd *io = (cd *)data + b * n;
                memcpy(x, io, sizeof(cd) * (size_t)n);
                cd *res;
                stockham_d(x, y, n, f, nf, w, &res);
                memcpy(io, res, sizeof(cd) * (size_t)n);
            }
        }
        free(x);
        free(y);
    }
    free(w);
    return err;
}

/* Strided FP64 C2C along one axis of a row-major complex array of shape [outer][n][inner]:
 * element (o, j, i) sits at ((o*n + j)*inner + i). Used for the 2D/3D separable passes. */
int oracle_c2c_axis_f64(double *data, int64_t outer, int64_t n, int64_t inner, int sign)
{
    if (n <= 0 || outer < 0 || inner < 0) return -1;
    if (n == 1) return 0;
    if (inner == 1) return oracle_c2c_f64(data, n, outer, sign);
    int64_t f[64];
    int nf = factorize(n, f);
    cd *w = (cd *)malloc(sizeof(cd) * (size_t)n);
    if (!w) return -2;
    roots_ld(w, n, sign);
    int err = 0;
#pragma omp parallel
    {
        cd *x = (cd *)malloc(sizeof(cd) * (size_t)n);
        cd *y = (cd *)malloc(sizeof(cd) * (size_t)n);
        if (!x || !y) {
#pragma omp atomic write
            err = -2;
        } else {
#pragma omp for schedule(static)
            for (int64_t t = 0; t < outer * inner; t++) {
                int64_t o = t / inner, i = t % inner;
                cd *base = (cd *)data + o * n * inner + i;
                for (int64_t j = 0; j < n; j++) x[j] = base[j * inner];
                cd *res;
                stockham_d(x, y, n, f, nf, w, &res);
                for (int64_t j = 0; j < n; j++) base[j * inner] = res[j];
            }
        }
        free(x);
        free(y);
    }
    free(w);
    return err;
}

/* R2C along the last axis: real in[batch][n] -> complex out[batch][n/2+1], sign by direction. */
int oracle_r2c_f64(const double *in, double *out, int64_t n, int64_t batch, int sign)
{
    if (n <= 0 || batch < 0) return -1;
    int64_t h = n / 2 + 1;
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)(n * (batch ? batch : 1)));
    if (!tmp) return -2;
    for (int64_t b = 0; b < batch; b++)
        for (int64_t j = 0; j < n; j++) {
            tmp[2 * (b * n + j)] = in[b * n + j];
            tmp[2 * (b * n + j) + 1] = 0.0;
        }
    int rc = oracle_c2c_f64(tmp, n, batch, sign);
    if (rc == 0)
        for (int64_t b = 0; b < batch; b++) memcpy(out + 2 * b * h, tmp + 2 * b * n, sizeof(double) * 2 * (size_t)h);
    free(tmp);
    return rc;
}

/* C2R along the last axis: complex in[batch][n/2+1] -> real out[batch][n]. The input is extended
 * Hermitian-wise (X[n-k] = conj X[k] for k > n/2, imaginary parts of X[0] and, for even n, X[n/2]
 * dropped), transformed with the given sign and the real part kept. */
int oracle_c2r_f64(const double *in, double *out, int64_t n, int64_t batch, int sign)
{
    if (n <= 0 || batch < 0) return -1;
    int64_t h = n / 2 + 1;
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)(n * (batch ? batch : 1)));
    if (!tmp) return -2;
    for (int64_t b = 0; b < batch; b++) {
        const double *src = in + 2 * b * h;
        double *x = tmp + 2 * b * n;
        for (int64_t k = 0; k < n; k++) {
            if (k < h) {
                x[2 * k] = src[2 * k];
                x[2 * k + 1] = src[2 * k + 1];
            } else {
                x[2 * k] = src[2 * (n - k)];
                x[2 * k + 1] = -src[2 * (n - k) + 1];
            }
        }
        x[1] = 0.0;
        if (n % 2 == 0) x[2 * (n / 2) + 1] = 0.0;
    }
    int rc = oracle_c2c_f64(tmp, n, batch, sign);
    if (rc == 0)
        for (int64_t b = 0; b < batch; b++)
            for (int64_t j = 0; j < n; j++) out[b * n + j] = tmp[2 * (b * n + j)];
    free(tmp);
    return rc;
}

/* ---------------------------------------------------------------- brute-force long double DFT */

/* O(n^2) DFT of one complex sequence in long double; roots indexed by (j*k mod n) so every factor
 * is one exact table entry. Used to cross-check the FFT above. */
int oracle_dft_ld(const double *in, double *out, int64_t n, int sign)
{
    if (n <= 0) return -1;
    const long double two_pi = 6.283185307179586476925286766559005768L;
    long double *wr = (long double *)malloc(sizeof(long double) * (size_t)n);
    long double *wi = (long double *)malloc(sizeof(long double) * (size_t)n);
    if (!wr || !wi) { free(wr); free(wi); return -2; }
    for (int64_t m = 0; m < n; m++) {
        long double a = two_pi * (long double)m / (long double)n;
        wr[m] = cosl(a);
        wi[m] = sign * sinl(a);
    }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
        long double re = 0.0L, im = 0.0L;
        for (int64_t j = 0; j < n; j++) {
            int64_t m = (int64_t)(((__int128)j * k) % n);
            long double xr = in[2 * j], xi = in[2 * j + 1];
            re += xr * wr[m] - xi * wi[m];
            im += xr * wi[m] + xi * wr[m];
        }
        out[2 * k] = (double)re;
        out[2 * k + 1] = (double)im;
    }
    free(wr);
    free(wi);
    return 0;
}

/* ---------------------------------------------------------------- error metrics */

/* rel-L2 = ||got - ref||_2 / ||ref||_2 and max-abs = max |got - ref| over `count` reals, where got
 * is FP32 and ref FP64 (the error report every parity check prints). */
void oracle_error_f32(const float *got, const double *ref, int64_t count, double *rel_l2, double *max_abs)
{
    double num = 0.0, den = 0.0, mx = 0.0;
#pragma omp parallel for reduction(+ : num, den) reduction(max : mx) schedule(static)
    for (int64_t i = 0; i < count; i++) {
        double d = (double)got[i] - ref[i];
        num += d * d;
        den += ref[i] * ref[i];
        double a = fabs(d);
        if (a > mx) mx = a;
    }
    *rel_l2 = den > 0.0 ? sqrt(num / den) : sqrt(num);
    *max_abs = mx;
}

/* ---------------------------------------------------------------- FP32 CPU baseline */

static void stockham_f(cf *x, cf *y, int64_t n, const int64_t *f, int nf, const cf *w, cf **result)
{
    int64_t ns = 1;
    cf *src = x, *dst = y;
    cf v[64], V[64];
    for (int s = 0; s < nf; s++) {
        int64_t R = f[s];
        int64_t m = n / R;
        int64_t tstep = n / (ns * R);
        for (int64_t j = 0; j < m; j++) {
            int64_t k = j % ns;
            for (int64_t r = 0; r < R; r++) {
                cf a = src[j + r * m];
                cf t = w[(k * r * tstep) % n];
                v[r].re = a.re * t.re - a.im * t.im;
                v[r].im = a.re * t.im + a.im * t.re;
            }
            for (int64_t q = 0; q < R; q++) {
                float re = 0.0f, im = 0.0f;
                for (int64_t r = 0; r < R; r++) {
                    cf t = w[((r * q) % R) * m];
                    re += v[r].re * t.re - v[r].im * t.im;
                    im += v[r].re * t.im + v[r].im * t.re;
                }
                V[q].re = re;
                V[q].im = im;
            }
            int64_t d = (j / ns) * ns * R + k;
            for (int64_t q = 0; q < R; q++) dst[d + q * ns] = V[q];
        }
        cf *t = src; src = dst; dst = t;
        ns *= R;
    }
    *result = src;
}

/* Batched FP32 C2C, out-of-place (in and out may alias), OpenMP over the batch with `threads`
 * threads (0 = runtime default). Radices are the prime factorisation with radix-4 first, at most
 * 64 per stage. Returns 0 on success, -3 when a prime factor exceeds 64. */
int oracle_c2c_f32(const float *in, float *out, int64_t n, int64_t batch, int sign, int threads)
{
    if (n <= 0 || batch < 0) return -1;
    int64_t f[64];
    int nf = factorize(n, f);
    for (int i = 0; i < nf; i++)
        if (f[i] > 64) return -3;
    cd *wd = (cd *)malloc(sizeof(cd) * (size_t)n);
    cf *w = (cf *)malloc(sizeof(cf) * (size_t)n);
    if (!wd || !w) { free(wd); free(w); return -2; }
    roots_ld(wd, n, sign);
    for (int64_t m = 0; m < n; m++) { w[m].re = (float)wd[m].re; w[m].im = (float)wd[m].im; }
    free(wd);
#ifdef _OPENMP
    int nth = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nth)
#endif
    {
        cf *x = (cf *)malloc(sizeof(cf) * (size_t)n);
        cf *y = (cf *)malloc(sizeof(cf) * (size_t)n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t b = 0; b < batch; b++) {
            memcpy(x, (const cf *)in + b * n, sizeof(cf) * (size_t)n);
            cf *res = x;
            if (n > 1) stockham_f(x, y, n, f, nf, w, &res);
            memcpy((cf *)out + b * n, res, sizeof(cf) * (size_t)n);
        }
        free(x);
        free(y);
    }
    free(w);
    return 0;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
